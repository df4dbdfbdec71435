// mmc_io.cpp -- ini parser and OVF reader/writer (host only).
//
// The reference's parse_input_ini (src/parse_input_ini.h:15-47) and
// ovf_io::parseOvfFile / writetoOvfFile (src/ovf_io.h:88-105) have empty
// bodies; this restates the formats they name.  Header lines follow the
// reference's parseHeaderLine (src/ovf_io.h:66-85): "# key: value # comment",
// key and value trimmed and lower-cased.
#include "mmc.h"
#include "mmc_io.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

void set_msg(char *msg, int len, const std::string &s)
{
    if (msg && len > 0) {
        std::snprintf(msg, (size_t)len, "%s", s.c_str());
    }
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// "# key: value # comment" -> (key, value), lower-cased (src/ovf_io.h:66-85)
bool header_line(const std::string &line, std::string &key, std::string &value)
{
    size_t colon = line.find(':');
    if (colon == std::string::npos) return false;
    std::string k = line.substr(0, colon);
    std::string v = line.substr(colon + 1);
    size_t hash = v.find('#');
    if (hash != std::string::npos) v = v.substr(0, hash);
    size_t p = 0;
    while (p < k.size() && (k[p] == '#' || std::isspace((unsigned char)k[p]))) ++p;
    key = lower(trim(k.substr(p)));
    value = lower(trim(v));
    return true;
}

bool little_endian_host()
{
    const uint16_t one = 1;
    return *reinterpret_cast<const uint8_t *>(&one) == 1;
}

template <typename T>
T swap_bytes(T v)
{
    unsigned char *p = reinterpret_cast<unsigned char *>(&v);
    std::reverse(p, p + sizeof(T));
    return v;
}

bool parse_double(const std::string &s, double &out)
{
    const char *b = s.c_str();
    char *e = nullptr;
    errno = 0;
    out = std::strtod(b, &e);
    if (e == b || errno == ERANGE) return false;
    while (*e && std::isspace((unsigned char)*e)) ++e;
    return *e == '\0';
}

bool parse_long(const std::string &s, long long &out)
{
    const char *b = s.c_str();
    char *e = nullptr;
    errno = 0;
    out = std::strtoll(b, &e, 10);
    if (e == b || errno == ERANGE) {
        // accept integral floating notation such as 1e5
        double d;
        if (!parse_double(s, d) || d != std::floor(d) || std::fabs(d) > 9.2e18) return false;
        out = (long long)d;
        return true;
    }
    while (*e && std::isspace((unsigned char)*e)) ++e;
    if (*e) {
        double d;
        if (!parse_double(s, d) || d != std::floor(d) || std::fabs(d) > 9.2e18) return false;
        out = (long long)d;
    }
    return true;
}

bool parse_bool(const std::string &s, int &out)
{
    std::string v = lower(s);
    if (v == "true" || v == "1") { out = 1; return true; }
    if (v == "false" || v == "0") { out = 0; return true; }
    return false;
}

}  // namespace

extern "C" {

int mmc_ovf_mesh_equal(const mmc_ovf_info *a, const mmc_ovf_info *b)
{
    auto close = [](double x, double y) { return std::fabs(x - y) <= 1e-5 * (std::fabs(x) + std::fabs(y)); };
    return a->xnodes == b->xnodes && a->ynodes == b->ynodes && a->znodes == b->znodes &&
           close(a->xstepsize, b->xstepsize) && close(a->ystepsize, b->ystepsize) &&
           close(a->zstepsize, b->zstepsize);
}

int mmc_ovf_read(const char *path, mmc_ovf_info *info, double *mx, double *my, double *mz,
                 long capacity, char *msg, int msglen)
{
    if (!path || !info) return MMC_EINVAL;
    std::ifstream f(path, std::ios::binary);
    if (!f) { set_msg(msg, msglen, std::string("cannot open ") + path); return MMC_EIO; }
    info->xnodes = info->ynodes = info->znodes = -1;
    info->xstepsize = info->ystepsize = info->zstepsize = -1;
    int version = 2, valuedim = 3;
    std::string line, key, value;
    std::string data_kind;
    bool first = true;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (first) {
            first = false;
            std::string l = lower(line);
            if (l.find("oommf") != std::string::npos && l.find("1.0") != std::string::npos) version = 1;
        }
        if (line.empty() || line[0] != '#') continue;
        if (!header_line(line, key, value)) continue;
        if (key == "xnodes") info->xnodes = std::atoi(value.c_str());
        else if (key == "ynodes") info->ynodes = std::atoi(value.c_str());
        else if (key == "znodes") info->znodes = std::atoi(value.c_str());
        else if (key == "xstepsize") parse_double(value, info->xstepsize);
        else if (key == "ystepsize") parse_double(value, info->ystepsize);
        else if (key == "zstepsize") parse_double(value, info->zstepsize);
        else if (key == "valuedim") valuedim = std::atoi(value.c_str());
        else if (key == "begin" && value.rfind("data", 0) == 0) { data_kind = trim(value.substr(4)); break; }
    }
    if (data_kind.empty()) { set_msg(msg, msglen, "no '# Begin: Data' block"); return MMC_EIO; }
    if (info->xnodes <= 0 || info->ynodes <= 0 || info->znodes <= 0) {
        set_msg(msg, msglen, "missing or invalid x/y/znodes");
        return MMC_EIO;
    }
    if (valuedim != 3) { set_msg(msg, msglen, "valuedim must be 3"); return MMC_EIO; }
    if (!mx || !my || !mz) return MMC_OK;
    const long N = (long)info->xnodes * info->ynodes * info->znodes;
    if (N > capacity) {
        set_msg(msg, msglen, "file holds " + std::to_string(N) + " cells, buffer " + std::to_string(capacity));
        return MMC_EINVAL;
    }
    if (data_kind == "text") {
        for (long i = 0; i < N; ++i) {
            double v[3];
            for (int k = 0; k < 3; ++k) {
                std::string tok;
                if (!(f >> tok)) { set_msg(msg, msglen, "truncated text data"); return MMC_EIO; }
                if (!parse_double(tok, v[k])) { set_msg(msg, msglen, "bad number '" + tok + "'"); return MMC_EIO; }
            }
            mx[i] = v[0]; my[i] = v[1]; mz[i] = v[2];
        }
        return MMC_OK;
    }
    const bool b4 = data_kind == "binary 4", b8 = data_kind == "binary 8";
    if (!b4 && !b8) { set_msg(msg, msglen, "unsupported data kind '" + data_kind + "'"); return MMC_EUNSUP; }
    // OVF 2.0 binary is little-endian, OVF 1.0 big-endian
    const bool swap = (version == 1) == little_endian_host();
    if (b4) {
        std::vector<float> buf(3 * N + 1);
        f.read(reinterpret_cast<char *>(buf.data()), sizeof(float) * buf.size());
        if (!f) { set_msg(msg, msglen, "truncated binary 4 data"); return MMC_EIO; }
        if (swap) for (auto &v : buf) v = swap_bytes(v);
        if (buf[0] != 1234567.0f) { set_msg(msg, msglen, "bad binary 4 check value"); return MMC_EIO; }
        for (long i = 0; i < N; ++i) { mx[i] = buf[1 + 3 * i]; my[i] = buf[2 + 3 * i]; mz[i] = buf[3 + 3 * i]; }
    } else {
        std::vector<double> buf(3 * N + 1);
        f.read(reinterpret_cast<char *>(buf.data()), sizeof(double) * buf.size());
        if (!f) { set_msg(msg, msglen, "truncated binary 8 data"); return MMC_EIO; }
        if (swap) for (auto &v : buf) v = swap_bytes(v);
        if (buf[0] != 123456789012345.0) { set_msg(msg, msglen, "bad binary 8 check value"); return MMC_EIO; }
        for (long i = 0; i < N; ++i) { mx[i] = buf[1 + 3 * i]; my[i] = buf[2 + 3 * i]; mz[i] = buf[3 + 3 * i]; }
    }
    return MMC_OK;
}

int mmc_ovf_write(const char *path, const mmc_ovf_info *in, const double *mx, const double *my,
                  const double *mz, int format, const char *title)
{
    if (!path || !in || !mx || !my || !mz) return MMC_EINVAL;
    if (format != 0 && format != 4 && format != 8) return MMC_EINVAL;
    if (!little_endian_host() && format != 0) return MMC_EUNSUP;
    std::ofstream f(path, std::ios::binary);
    if (!f) return MMC_EIO;
    const long N = (long)in->xnodes * in->ynodes * in->znodes;
    const char *kind = format == 0 ? "Text" : (format == 4 ? "Binary 4" : "Binary 8");
    char buf[256];
    f << "# OOMMF OVF 2.0\n# Segment count: 1\n# Begin: Segment\n# Begin: Header\n";
    f << "# Title: " << (title ? title : "m") << "\n# meshtype: rectangular\n# meshunit: m\n";
    std::snprintf(buf, sizeof buf, "# xmin: 0\n# ymin: 0\n# zmin: 0\n# xmax: %.17g\n# ymax: %.17g\n# zmax: %.17g\n",
                  in->xnodes * in->xstepsize, in->ynodes * in->ystepsize, in->znodes * in->zstepsize);
    f << buf;
    f << "# valuedim: 3\n# valuelabels: m_x m_y m_z\n# valueunits: 1 1 1\n";
    std::snprintf(buf, sizeof buf, "# xbase: %.17g\n# ybase: %.17g\n# zbase: %.17g\n",
                  in->xstepsize / 2, in->ystepsize / 2, in->zstepsize / 2);
    f << buf;
    f << "# xnodes: " << in->xnodes << "\n# ynodes: " << in->ynodes << "\n# znodes: " << in->znodes << "\n";
    std::snprintf(buf, sizeof buf, "# xstepsize: %.17g\n# ystepsize: %.17g\n# zstepsize: %.17g\n",
                  in->xstepsize, in->ystepsize, in->zstepsize);
    f << buf;
    f << "# End: Header\n# Begin: Data " << kind << "\n";
    if (format == 0) {
        for (long i = 0; i < N; ++i) {
            std::snprintf(buf, sizeof buf, "%.17g %.17g %.17g\n", mx[i], my[i], mz[i]);
            f << buf;
        }
    } else if (format == 4) {
        std::vector<float> out(3 * N + 1);
        out[0] = 1234567.0f;
        for (long i = 0; i < N; ++i) { out[1 + 3 * i] = (float)mx[i]; out[2 + 3 * i] = (float)my[i]; out[3 + 3 * i] = (float)mz[i]; }
        f.write(reinterpret_cast<const char *>(out.data()), sizeof(float) * out.size());
        f << "\n";
    } else {
        std::vector<double> out(3 * N + 1);
        out[0] = 123456789012345.0;
        for (long i = 0; i < N; ++i) { out[1 + 3 * i] = mx[i]; out[2 + 3 * i] = my[i]; out[3 + 3 * i] = mz[i]; }
        f.write(reinterpret_cast<const char *>(out.data()), sizeof(double) * out.size());
        f << "\n";
    }
    f << "# End: Data " << kind << "\n# End: Segment\n";
    return f ? MMC_OK : MMC_EIO;
}

int mmc_ini_parse(const char *path, mmc_config *cfg, char *msg, int msglen)
{
    if (!path || !cfg) return MMC_EINVAL;
    std::ifstream f(path);
    if (!f) { set_msg(msg, msglen, std::string("cannot open ") + path); return MMC_EIO; }
    std::map<std::string, std::string> kv;
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#') continue;
        if (t[0] == '[') continue;                      // sections are ignored
        size_t eq = t.find('=');
        if (eq == std::string::npos) {
            set_msg(msg, msglen, std::string(path) + ":" + std::to_string(lineno) + ": expected key = value");
            return MMC_EINVAL;
        }
        std::string k = trim(t.substr(0, eq)), v = trim(t.substr(eq + 1));
        // trailing " ; comment" / " # comment"
        for (const char *c : {" ;", "\t;", " #", "\t#"}) {
            size_t p = v.find(c);
            if (p != std::string::npos) v = trim(v.substr(0, p));
        }
        if (k.empty()) {
            set_msg(msg, msglen, std::string(path) + ":" + std::to_string(lineno) + ": empty key");
            return MMC_EINVAL;
        }
        kv[k] = v;
    }
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->precision = MMC_FP32;
    cfg->ovf_format = 8;
    std::string err;
    auto need = [&](const char *k) -> const std::string * {
        auto it = kv.find(k);
        if (it == kv.end()) { if (err.empty()) err = std::string("missing key '") + k + "'"; return nullptr; }
        return &it->second;
    };
    auto get_d = [&](const char *k, double &out) {
        const std::string *v = need(k);
        if (v && !parse_double(*v, out) && err.empty()) err = std::string("key '") + k + "': not a number";
    };
    auto get_i = [&](const char *k, int &out) {
        const std::string *v = need(k);
        long long t = 0;
        if (v && !parse_long(*v, t) && err.empty()) err = std::string("key '") + k + "': not an integer";
        out = (int)t;
    };
    auto get_l = [&](const char *k, long &out) {
        const std::string *v = need(k);
        long long t = 0;
        if (v && !parse_long(*v, t) && err.empty()) err = std::string("key '") + k + "': not an integer";
        out = (long)t;
    };
    auto get_b = [&](const char *k, int &out) {
        const std::string *v = need(k);
        if (v && !parse_bool(*v, out) && err.empty()) err = std::string("key '") + k + "': not a bool";
    };
    get_i("nx", cfg->nx); get_i("ny", cfg->ny); get_i("nz", cfg->nz);
    get_d("dx", cfg->dx); get_d("dy", cfg->dy); get_d("dz", cfg->dz);
    get_d("Bextx", cfg->Bextx); get_d("Bexty", cfg->Bexty); get_d("Bextz", cfg->Bextz);
    get_i("pbc_x", cfg->pbc_x); get_i("pbc_y", cfg->pbc_y); get_i("pbc_z", cfg->pbc_z);
    get_d("ms", cfg->ms); get_d("Aex", cfg->Aex); get_d("Dind", cfg->Dind);
    get_d("anisUx", cfg->anisUx); get_d("anisUy", cfg->anisUy); get_d("anisUz", cfg->anisUz);
    get_d("Ku1", cfg->Ku1);
    get_b("use_random_init", cfg->use_random_init);
    {
        const std::string *v = need("randomseed");
        long long t = 0;
        if (v && !parse_long(*v, t) && err.empty()) err = "key 'randomseed': not an integer";
        cfg->randomseed = t;
    }
    {
        auto it = kv.find("ovf_filename");
        if (it != kv.end()) {
            std::string v = it->second;
            if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
            std::snprintf(cfg->ovf_filename, sizeof cfg->ovf_filename, "%s", v.c_str());
        } else if (!cfg->use_random_init && err.empty()) {
            err = "missing key 'ovf_filename' (use_random_init = false)";
        }
    }
    get_d("Temperature_start", cfg->Temperature_start);
    get_d("Temperature_end", cfg->Temperature_end);
    get_d("Temperature_step", cfg->Temperature_step);
    get_b("Temperature_use_exp", cfg->Temperature_use_exp);
    get_l("circle_per_stage", cfg->circle_per_stage);
    get_l("terminal_output_period", cfg->terminal_output_period);
    get_l("energy_output_period", cfg->energy_output_period);
    get_l("ms_output_period", cfg->ms_output_period);
    get_i("rand_block_size", cfg->rand_block_size);
    get_b("cal_demag_flag", cfg->cal_demag_flag);
    {
        auto it = kv.find("precision");
        if (it != kv.end()) {
            std::string v = lower(it->second);
            if (v == "fp32" || v == "float" || v == "single") cfg->precision = MMC_FP32;
            else if (v == "fp64" || v == "double") cfg->precision = MMC_FP64;
            else if (err.empty()) err = "key 'precision': expected fp32 or fp64";
        }
        it = kv.find("ovf_format");
        if (it != kv.end()) {
            long long t = -1;
            if (!parse_long(it->second, t) || (t != 0 && t != 4 && t != 8)) {
                if (err.empty()) err = "key 'ovf_format': expected 0, 4 or 8";
            } else cfg->ovf_format = (int)t;
        }
    }
    if (err.empty()) {
        if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) err = "nx, ny, nz must be >= 1";
        else if (cfg->circle_per_stage < 0) err = "circle_per_stage must be >= 0";
        else if (!(cfg->Temperature_start > 0)) err = "Temperature_start must be > 0";
        else if (cfg->Temperature_use_exp && !(cfg->Temperature_step > 0 && cfg->Temperature_step < 1))
            err = "Temperature_step must be in (0,1) with Temperature_use_exp";
        else if (!cfg->Temperature_use_exp && !(cfg->Temperature_step > 0))
            err = "Temperature_step must be > 0";
    }
    if (!err.empty()) { set_msg(msg, msglen, err); return MMC_EINVAL; }
    return MMC_OK;
}

}  // extern "C"
