// Reader of the reference's block-structured ".env" configuration files
// (SDK/sutil/Environment.cpp:593-792) for the options the MDAS apps register
// (SDK/optixMotionBlur/optixMotionBlur.cpp:107-134 and the other four apps),
// plus a `Scene` block choosing the analytic integrand.
//
// Format: `Block { key value ... }`, '#' starts a comment, tokens are
// whitespace separated, several options may share a line.  Option types and
// value counts follow the reference (int, float, bool, 3-vector, 4-vector,
// 11-value frame, 9-value parallelogram, string).  The first occurrence of an
// option wins when it is read back (Environment.cpp:143-159).  Unknown
// options, unpaired braces and values that do not parse are errors naming the
// file and line (the reference prints and returns false).
#include <ctype.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mdas.h"

namespace {

enum OptType { O_INT, O_FLOAT, O_BOOL, O_VEC3, O_VEC4, O_FRAME, O_PARA, O_STRING };

struct OptDef {
    const char* name;
    OptType type;
    const char* def;
};

// Union of the options the five MDAS apps register, with their defaults.
const OptDef kOptions[] = {
    {"Camera.position", O_VEC3, nullptr},
    {"Camera.direction", O_VEC3, nullptr},
    {"Camera.upVector", O_VEC3, "0.0 1.0 0.0"},
    {"Camera.fovy", O_FLOAT, "45.0"},
    {"Camera.focalDistance", O_FLOAT, "0.0"},
    {"Camera.lensRadius", O_FLOAT, "0.0"},
    {"Film.filename", O_STRING, nullptr},
    {"Film.width", O_INT, "1024"},
    {"Film.height", O_INT, "768"},
    {"Sampler.samples", O_FLOAT, "1"},
    {"Sampler.mdas", O_BOOL, "0"},
    {"Mdas.scaleFactor", O_FLOAT, "1.0"},
    {"Mdas.alpha", O_FLOAT, "0.25"},
    {"Mdas.bitsPerDim", O_INT, "0"},
    {"Mdas.extraImgBits", O_INT, "8"},
    {"Mdas.candidatesNum", O_INT, "4"},
    {"Light.point", O_VEC3, nullptr},
    {"Light.distant", O_VEC3, nullptr},
    {"Light.color", O_VEC3, nullptr},
    {"Light.area", O_PARA, nullptr},
    {"Light.background", O_VEC3, "1.0 1.0 1.0"},
    {"Model.filename", O_STRING, nullptr},
    {"Model.frame", O_FRAME, nullptr},
    {"Model.emissive", O_BOOL, "0"},
    {"EnvironmentMap.filename", O_STRING, nullptr},
    // additions: the analytic integrand
    {"Scene.kind", O_STRING, nullptr},
    {"Scene.seed", O_INT, "0"},
    {"Scene.dim", O_INT, "0"},
};

int value_count(OptType t) {
    switch (t) {
    case O_VEC3: return 3;
    case O_VEC4: return 4;
    case O_FRAME: return 11;
    case O_PARA: return 9;
    default: return 1;
    }
}

// A number token parses when it starts with a number (like stream extraction).
bool number_prefix(const std::string& s, bool integer) {
    const char* p = s.c_str();
    char* end = nullptr;
    if (integer) (void)strtol(p, &end, 10);
    else (void)strtof(p, &end);
    return end != p;
}

bool parse_bool(const std::string& v, bool* out) {
    static const char* yes[] = {"true", "yes", "on", "1", "TRUE", "YES", "ON"};
    static const char* no[] = {"false", "no", "off", "0", "FALSE", "NO", "OFF"};
    for (const char* y : yes)
        if (v == y) { *out = true; return true; }
    for (const char* n : no)
        if (v == n) { *out = false; return true; }
    return false;
}

const OptDef* find(const std::string& name) {
    for (const OptDef& o : kOptions)
        if (name == o.name) return &o;
    return nullptr;
}

struct Values {
    std::map<std::string, std::vector<std::string>> v;  // option -> values (joined tokens)
    std::string first(const std::string& name) const {
        auto it = v.find(name);
        if (it != v.end() && !it->second.empty()) return it->second.front();
        const OptDef* d = find(name);
        return d && d->def ? d->def : "";
    }
};

thread_local std::string t_env_err;

int env_fail(const std::string& m) {
    t_env_err = m;
    return MDAS_ERR_INVALID;
}

int read_env(const char* path, Values& out) {
    std::ifstream in(path);
    if (!in.good()) return env_fail(std::string("cannot open '") + path + "'");
    std::string line, block, last;
    bool inblock = false;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        size_t c = line.find('#');
        if (c != std::string::npos) line.resize(c);
        std::istringstream ss(line);
        std::vector<std::string> w;
        for (std::string t; ss >> t;) w.push_back(t);
        for (size_t i = 0; i < w.size(); ++i) {
            const std::string where = std::string(" in '") + path + "' (line " + std::to_string(lineno) + ")";
            if (w[i] == "{") {
                if (last.empty()) return env_fail("block without a name" + where);
                block = last;
                inblock = true;
            } else if (w[i] == "}") {
                if (!inblock) return env_fail("unpaired }" + where);
                inblock = false;
                block.clear();
            } else if (inblock) {
                const std::string name = block + "." + w[i];
                const OptDef* d = find(name);
                if (!d) return env_fail("unknown option '" + name + "'" + where);
                const int n = value_count(d->type);
                if (i + n >= w.size()) return env_fail("missing value for '" + name + "'" + where);
                std::string joined;
                for (int k = 1; k <= n; ++k) {
                    const std::string& tok = w[i + k];
                    bool ok = true;
                    if (d->type == O_INT) ok = number_prefix(tok, true);
                    else if (d->type == O_BOOL) {
                        bool b;
                        ok = parse_bool(tok, &b);
                    } else if (d->type != O_STRING) ok = number_prefix(tok, false);
                    if (!ok) return env_fail("bad value '" + tok + "' for '" + name + "'" + where);
                    if (k > 1) joined += " ";
                    joined += tok;
                }
                out.v[name].push_back(joined);
                i += n;
            }
            last = w[i];
        }
    }
    if (inblock) return env_fail(std::string("unterminated block '") + block + "' in '" + path + "'");
    return MDAS_OK;
}

int scene_kind(const std::string& k) {
    if (k == "constant") return MDAS_SCENE_CONSTANT;
    if (k == "disks" || k == "disks_mb") return MDAS_SCENE_DISKS_MB;
    if (k == "dof" || k == "spheres_dof") return MDAS_SCENE_SPHERES_DOF;
    if (k == "mbdof" || k == "spheres_mbdof") return MDAS_SCENE_SPHERES_MBDOF;
    if (k == "ao") return MDAS_SCENE_AO;
    if (k == "ramp") return MDAS_SCENE_RAMP;
    return -1;
}

}  // namespace

extern "C" {

const char* mdas_env_last_error(void) { return t_env_err.c_str(); }

int mdas_env_read(const char* path, mdas_env* out) {
    if (!path || !out) return env_fail("null argument");
    memset(out, 0, sizeof(*out));
    Values v;
    int rc = read_env(path, v);
    if (rc) return rc;
    bool b = false;
    parse_bool(v.first("Sampler.mdas"), &b);
    out->mdas = b ? 1 : 0;
    out->samples = strtof(v.first("Sampler.samples").c_str(), nullptr);
    out->scale_factor = strtof(v.first("Mdas.scaleFactor").c_str(), nullptr);
    out->alpha = strtof(v.first("Mdas.alpha").c_str(), nullptr);
    out->bits_per_dim = (int)strtol(v.first("Mdas.bitsPerDim").c_str(), nullptr, 10);
    out->extra_img_bits = (int)strtol(v.first("Mdas.extraImgBits").c_str(), nullptr, 10);
    out->candidates_num = (int)strtol(v.first("Mdas.candidatesNum").c_str(), nullptr, 10);
    out->width = (int)strtol(v.first("Film.width").c_str(), nullptr, 10);
    out->height = (int)strtol(v.first("Film.height").c_str(), nullptr, 10);
    std::string fn = v.first("Film.filename");
    strncpy(out->filename, fn.c_str(), sizeof(out->filename) - 1);
    std::string kind = v.first("Scene.kind");
    out->scene_kind = kind.empty() ? -1 : scene_kind(kind);
    if (!kind.empty() && out->scene_kind < 0) return env_fail("unknown Scene.kind '" + kind + "'");
    out->scene_seed = (uint32_t)strtoul(v.first("Scene.seed").c_str(), nullptr, 10);
    out->dim = (int)strtol(v.first("Scene.dim").c_str(), nullptr, 10);
    if (out->dim == 0) {
        static const int dims[6] = {3, 3, 4, 5, 6, 3};
        out->dim = out->scene_kind >= 0 ? dims[out->scene_kind] : 3;
    }
    return MDAS_OK;
}

int mdas_env_config(const mdas_env* env, long long max_samples, mdas_config* cfg) {
    if (!env || !cfg) return env_fail("null argument");
    memset(cfg, 0, sizeof(*cfg));
    cfg->dim = env->dim;
    cfg->max_samples = max_samples;
    cfg->candidates_num = env->candidates_num;
    cfg->bits_per_dim = env->bits_per_dim;
    cfg->extra_img_bits = env->extra_img_bits;
    cfg->alpha = env->alpha;
    // optixMotionBlur.cpp:347-348: scaleFactor * width, scaleFactor * height
    cfg->scale_x = env->scale_factor * (float)env->width;
    cfg->scale_y = env->scale_factor * (float)env->height;
    cfg->log_path = nullptr;
    return MDAS_OK;
}

}  // extern "C"
