// sph_xml.cpp - host-only scene-file front end of the C ABI: the scene list
// of an XML scene file, each scene's values, and the particle fill of a
// scene's init volume.  Needs no device.
//
//   sph_xml_scene_count / sph_xml_scene ... cSPH::LoadScenes (source/SPH/SPH_Scenes.cpp:53-83),
//                                          Scene::InitDefault + Scene::_FromXML
//                                          (source/SPH/Scene.cpp:7-66, source/SPH/Scene_Load.cpp:11-111),
//                                          Scene::_UpdatePar/_UpdateGrid (source/SPH/Scene.cpp:76-113)
//   sph_scene_fill ........................ cSPH::Reset (source/SPH/SPH_Init.cpp:23-76)
//
// The reader keeps the tree a lenient XML reader builds from the reference's
// own files: an element left open absorbs the elements after it as children
// until the reader meets an end tag it does not expect; reading stops there
// and everything read so far is kept.  The scene list is the root element's
// <Scene> children in file order.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pibiti_sph.h"

namespace {

// ------------------------------------------------------------------ reader
struct Node {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<Node> kids;  // elements only

    const char* attr(const char* key) const {
        for (const auto& a : attrs)
            if (a.first == key) return a.second.c_str();
        return nullptr;
    }
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
inline bool name_start(unsigned char c) { return c >= 127 || std::isalpha(c) || c == '_'; }
inline bool name_char(unsigned char c) {
    return c >= 127 || std::isalnum(c) || c == '_' || c == '-' || c == '.' || c == ':';
}

class Reader {
public:
    explicit Reader(std::string text) : s_(std::move(text)) {}

    // Parse the document; returns its element children (possibly partial).
    std::vector<Node> document() {
        std::vector<Node> top;
        size_t i = skip(0);
        while (i < s_.size()) {
            if (s_[i] != '<') break;  // text at document level ends the document
            if (starts(i, "<!--")) {
                if (!comment(i)) break;
            } else if (starts(i, "<!") || starts(i, "<?")) {
                if (!unknown(i)) break;
            } else if (i + 1 < s_.size() && name_start((unsigned char)s_[i + 1])) {
                top.emplace_back();
                if (!element(i, top.back())) break;
            } else {
                if (!unknown(i)) break;
            }
            i = skip(i);
        }
        return top;
    }

private:
    std::string s_;

    size_t skip(size_t i) const {
        while (i < s_.size() && is_space(s_[i])) ++i;
        return i;
    }
    bool starts(size_t i, const char* lit) const { return s_.compare(i, std::strlen(lit), lit) == 0; }

    bool comment(size_t& i) {
        const size_t e = s_.find("-->", i + 4);
        if (e == std::string::npos) return false;
        i = e + 3;
        return true;
    }
    bool unknown(size_t& i) {
        const size_t e = s_.find('>', i + 1);
        if (e == std::string::npos) return false;
        i = e + 1;
        return true;
    }
    size_t name(size_t i, std::string& out) const {
        if (i >= s_.size() || !name_start((unsigned char)s_[i])) return std::string::npos;
        size_t j = i;
        while (j < s_.size() && name_char((unsigned char)s_[j])) ++j;
        out.assign(s_, i, j - i);
        return j;
    }
    // Attribute value text: the five predefined entities and character
    // references are decoded, anything else is kept as written.
    static void decode(const std::string& raw, std::string& out) {
        out.clear();
        for (size_t k = 0; k < raw.size(); ++k) {
            if (raw[k] != '&') {
                out += raw[k];
                continue;
            }
            static const struct { const char* ent; char c; } named[] = {
                {"&amp;", '&'}, {"&lt;", '<'}, {"&gt;", '>'}, {"&quot;", '"'}, {"&apos;", '\''}};
            bool done = false;
            for (const auto& e : named)
                if (raw.compare(k, std::strlen(e.ent), e.ent) == 0) {
                    out += e.c;
                    k += std::strlen(e.ent) - 1;
                    done = true;
                    break;
                }
            if (!done && k + 2 < raw.size() && raw[k + 1] == '#') {
                const size_t semi = raw.find(';', k);
                if (semi != std::string::npos) {
                    const bool hex = raw[k + 2] == 'x';
                    const long v = std::strtol(raw.c_str() + k + (hex ? 3 : 2), nullptr, hex ? 16 : 10);
                    if (v > 0 && v < 128) {
                        out += (char)v;
                        k = semi;
                        done = true;
                    }
                }
            }
            if (!done) out += '&';
        }
    }
    bool attribute(size_t& i, Node& n) {
        std::string key;
        size_t j = name(i, key);
        if (j == std::string::npos) return false;
        j = skip(j);
        if (j >= s_.size() || s_[j] != '=') return false;
        j = skip(j + 1);
        if (j >= s_.size()) return false;
        std::string raw;
        if (s_[j] == '"' || s_[j] == '\'') {
            const size_t e = s_.find(s_[j], j + 1);
            if (e == std::string::npos) return false;
            raw.assign(s_, j + 1, e - j - 1);
            j = e + 1;
        } else {
            while (j < s_.size() && !is_space(s_[j]) && s_[j] != '/' && s_[j] != '>') {
                if (s_[j] == '"' || s_[j] == '\'') return false;
                raw += s_[j++];
            }
        }
        if (n.attr(key.c_str())) return false;  // a repeated attribute is an error
        std::string val;
        decode(raw, val);
        n.attrs.emplace_back(std::move(key), std::move(val));
        i = j;
        return true;
    }
    // <name attrs> children </name>  or  <name attrs/>.  On an error the node
    // keeps what was read and false is returned.
    bool element(size_t& i, Node& n) {
        size_t j = name(skip(i + 1), n.name);
        if (j == std::string::npos) return false;
        for (;;) {
            j = skip(j);
            if (j >= s_.size()) return false;
            if (s_[j] == '/') {
                if (j + 1 >= s_.size() || s_[j + 1] != '>') return false;
                i = j + 2;
                return true;
            }
            if (s_[j] == '>') {
                ++j;
                break;
            }
            if (!attribute(j, n)) return false;
        }
        // children until an end tag
        for (;;) {
            j = skip(j);
            if (j >= s_.size()) return false;
            if (s_[j] != '<') {
                const size_t e = s_.find('<', j);
                if (e == std::string::npos) return false;
                j = e;
                continue;
            }
            if (starts(j, "</")) break;
            bool ok;
            if (starts(j, "<!--")) {
                ok = comment(j);
            } else if (starts(j, "<!") || starts(j, "<?") || j + 1 >= s_.size() ||
                       !name_start((unsigned char)s_[j + 1])) {
                ok = unknown(j);
            } else {
                n.kids.emplace_back();
                ok = element(j, n.kids.back());
            }
            if (!ok) return false;
        }
        const std::string end = "</" + n.name;
        if (!starts(j, end.c_str())) return false;
        j = skip(j + end.size());
        if (j >= s_.size() || s_[j] != '>') return false;
        i = j + 1;
        return true;
    }
};

std::vector<const Node*> scene_nodes(const std::vector<Node>& top) {
    std::vector<const Node*> out;
    const Node* root = top.empty() ? nullptr : &top.front();
    if (!root) return out;
    for (const auto& k : root->kids)
        if (k.name == "Scene") out.push_back(&k);
    return out;
}

// ------------------------------------------------------- value conversion
// The reference's toFloat ((float)atof), toInt (strtol base 0) and toVec3
// (sscanf "%f %f %f"; components the text does not give are 0 here, where
// the reference leaves them undefined).
float to_float(const char* a) { return (float)std::atof(a); }
int to_int(const char* a) { return (int)std::strtol(a, nullptr, 0); }
void to_vec3(const char* a, float v[3]) {
    float t[3] = {0.0f, 0.0f, 0.0f};
    if (std::sscanf(a, "%f %f %f", &t[0], &t[1], &t[2]) < 0) t[0] = t[1] = t[2] = 0.0f;
    for (int c = 0; c < 3; ++c) v[c] = t[c];
}

constexpr float kPi = 3.141592654f;  // source/CUDA/Params.cuh:18

// The values of class Scene (source/SPH/Scene.h:23-60) that the step, the
// init volume and the app's per-frame update read.
struct SceneVals {
    sph_scene s{};
    sph_scene_info info{};
    float coll_pos[3] = {-77.0f, 0.0f, 0.0f};
};

void defaults(SceneVals& v) {
    sph_scene_default(&v.s);
    sph_scene_info& in = v.info;
    std::memset(&in, 0, sizeof(in));
    std::strncpy(in.title, "No name", sizeof(in.title) - 1);
    const float w[3] = {0.2f, 0.25f, 0.2f};
    for (int c = 0; c < 3; ++c) {
        in.init_min[c] = -w[c];
        in.init_max[c] = w[c];
    }
    in.spacing = 1.38f;
    in.drop_r = 5.0f;
    in.rain = 0;
    in.init_type = 0;
    in.init_last = 1;
    in.cam_pos[1] = 0.04f;
    in.cam_pos[2] = -0.44f;
    in.cam_rot[0] = 5.0f;
    for (auto& e : in.emit) e.vel = 1.5f;  // Emitter() (source/SPH/Scene.h:12)
}

// Scene::_FromXML in the reference's statement order: where two attributes
// write the same value the later statement wins.
void from_xml(const Node& n, SceneVals& v) {
    sph_scene& s = v.s;
    sph_scene_info& in = v.info;
    const char* a;
#define A(key) if ((a = n.attr(key)))
    // world and init volume (Scene_Load.cpp:22-33)
    A("World") {
        float w[3] = {0.0f, 0.0f, 0.0f};
        to_vec3(a, w);
        for (int c = 0; c < 3; ++c) {
            s.world_min[c] = w[c] * -0.5f;
            s.world_max[c] = w[c] * 0.5f;
            in.init_min[c] = w[c] * -0.5f;
            in.init_max[c] = w[c] * 0.5f;
        }
    }
    A("Init") {
        float w[3] = {0.0f, 0.0f, 0.0f};
        to_vec3(a, w);
        for (int c = 0; c < 3; ++c) {
            in.init_min[c] = w[c] * -0.5f;
            in.init_max[c] = w[c] * 0.5f;
        }
    }
    A("WorldMin") to_vec3(a, s.world_min);
    A("InitMin") to_vec3(a, in.init_min);
    A("WorldMax") to_vec3(a, s.world_max);
    A("InitMax") to_vec3(a, in.init_max);
    A("InitType") in.init_type = to_int(a);
    A("InitLast") in.init_last = to_int(a);
    A("CellSize") s.cell_size = to_float(a);
    // camera, drops (:37-44)
    A("name") std::strncpy(in.title, a, sizeof(in.title) - 1);
    A("chapter") in.chapter = 1;
    A("CamPos") to_vec3(a, in.cam_pos);
    A("CamRot") to_vec3(a, in.cam_rot);
    A("dropR") in.drop_r = to_float(a);
    A("rain") in.rain = to_int(a);
    // emitters: one in attributes, up to four as children (:46-54)
    A("EmitSize") in.emit[0].size = to_int(a);
    A("EmitPos") to_vec3(a, in.emit[0].pos);
    A("EmitVel") in.emit[0].vel = to_float(a);
    A("EmitSize2") in.emit[0].size2 = to_int(a);
    A("EmitRot") to_vec3(a, in.emit[0].rot);
#undef A
    {
        int i = 0;
        for (const auto& k : n.kids) {
            if (k.name != "Emitter" || i >= 4) continue;
            sph_emitter& e = in.emit[i++];
            if ((a = k.attr("size"))) e.size = to_int(a);
            if ((a = k.attr("pos"))) to_vec3(a, e.pos);
            if ((a = k.attr("vel"))) e.vel = to_float(a);
            if ((a = k.attr("size2"))) e.size2 = to_int(a);
            if ((a = k.attr("rot"))) to_vec3(a, e.rot);
        }
        // accelerators as children (:57-62)
        i = 0;
        for (const auto& k : n.kids) {
            if (k.name != "Accel" || i >= SPH_NUM_ACCEL) continue;
            sph_accel& c = s.acc[i++];
            if ((a = k.attr("type"))) c.type = (uint32_t)to_int(a);
            if ((a = k.attr("pos"))) to_vec3(a, c.pos);
            if ((a = k.attr("size"))) to_vec3(a, c.size);
            if ((a = k.attr("acc"))) to_vec3(a, c.acc);
        }
    }
#define A(key) if ((a = n.attr(key)))
    // collider (:65-67)
    A("ColliderR") s.coll_r = to_float(a);
    A("ColliderPos") to_vec3(a, v.coll_pos);
    A("spring") s.coll_spring = to_float(a);
    A("damping") s.coll_damping = to_float(a);
    A("shear") s.coll_shear = to_float(a);
    // simulation (:71-74)
    A("Particles") s.num_particles = (uint32_t)to_int(a);
    A("ParticlesK") s.num_particles = (uint32_t)(to_int(a) * 1024);
    A("maxParInCell") s.max_par_in_cell = (uint32_t)to_int(a);
    A("TimeStep") s.time_step = to_float(a);
    A("globalDamping") s.global_damping = to_float(a);
    A("Gravity") to_vec3(a, s.gravity);
    // SPH (:77-81)
    A("particleR") s.particle_r = to_float(a);
    A("minDist") s.min_dist = to_float(a);
    A("particleH") s.h = to_float(a);
    A("spacing") in.spacing = to_float(a);
    A("RestDensity") s.rest_density = to_float(a);
    A("minDens") s.min_dens = to_float(a);
    A("Stiffness") s.stiffness = to_float(a);
    A("Viscosity") s.viscosity = to_float(a);
    // boundary (:84-88)
    A("distBndSoft") s.dist_bnd_soft = to_float(a);
    A("bndStiff") s.bnd_stiff = to_float(a);
    A("distBndHard") s.dist_bnd_hard = to_float(a);
    A("bndDamp") s.bnd_damp = to_float(a);
    A("bndDampC") s.bnd_damp_c = to_float(a);
    A("bndType") s.bnd_type = (uint32_t)to_int(a);
    A("bndEffZ") s.bnd_eff_z = (uint32_t)to_int(a);
    A("HmapType") s.hmap_type = to_int(a);
    // pump (:92-94)
    A("PumpAngOut") s.pump_ang_out = to_float(a);
    A("PumpHClose") s.h_close = to_float(a);
    A("PumpRadIn") s.pump_rad_in = to_float(a);
    A("ExitVel") s.exit_vel = to_float(a);
    A("ExitDist") s.exit_dist = to_float(a);
    A("s1") s.shape_s[0] = to_float(a);
    A("s2") s.shape_s[1] = to_float(a);
    A("s3") s.shape_s[2] = to_float(a);
    A("s4") s.shape_s[3] = to_float(a);
    A("s5") s.shape_s[4] = to_float(a);
    A("s6") s.shape_s[5] = to_float(a);
    // rotor (:96-103)
    A("RotorSizes") {
        float r3[3];
        to_vec3(a, r3);
        for (int c = 0; c < 3; ++c) s.rot_size[c] = (int32_t)r3[c];
    }
    A("RotorType") s.rot_type = to_int(a);
    A("Rotor2Dist") s.rot2_dist = to_float(a);
    A("RotorVel") s.rot_vel = to_float(a);
    A("Rotor2Vel") s.rot2_vel = to_float(a);
    A("RotorAngle") s.rot_angle = to_float(a) * kPi / 180.0f;
    A("Rotor2Angle") s.rot2_angle = to_float(a) * kPi / 180.0f;
    A("colParR") s.rot_r = to_float(a);
    A("colParSpc") s.rot_spc = to_float(a);
    A("RotorBlades") s.rot_blades = to_int(a);
    A("RotorTwist") s.rot_twist = to_float(a);
    // waves, aliases of rotor values (:106-108)
    A("WaveSpeed") {
        s.rot_vel = to_float(a);
        s.rot_angle = -kPi / 2.0f;
    }
    A("WaveAmpl") s.rot_twist = to_float(a);
    A("HSlope") s.rot2_angle = to_float(a);
    // height map, aliases again (:111-113)
    A("Hheight") s.rot2_angle = to_float(a);
    A("Hscale") s.h_close = to_float(a);
    A("HxFq") s.shape_s[0] = to_float(a);
    A("HxOfs") s.shape_s[1] = to_float(a);
    A("HzFq") s.shape_s[2] = to_float(a);
    A("HzOfs") s.shape_s[3] = to_float(a);
    A("HholeR") s.shape_s[4] = to_float(a);
#undef A
}

// Scene values -> sph_scene and the info record (the parts of
// Scene::_UpdatePar/_UpdateGrid that are not kernel constants).
void finish(SceneVals& v) {
    sph_scene& s = v.s;
    sph_scene_info& in = v.info;
    // the collider acts wherever no rotor runs (the force update decides)
    s.collider = 1;
    // the app's colour mode and dye clear count (source/SPH/Scene.cpp:61-63)
    s.clr_type = SPH_CLR_VEL_ACC;
    s.dye_clear = 2;
    for (int c = 0; c < 3; ++c) s.coll_pos[c] = v.coll_pos[c];
    // _UpdatePar: spacing *= particleR; _UpdateGrid: the init volume moves in
    // by distBndSoft*particleR - particleR on every side (Scene.cpp:80,100-102)
    const float r = s.particle_r;
    in.spacing = in.spacing * r;
    const float b = s.dist_bnd_soft * r - r;
    for (int c = 0; c < 3; ++c) {
        in.init_min[c] = in.init_min[c] + b;
        in.init_max[c] = in.init_max[c] - b;
    }
}

int read_scenes(const char* text, size_t len, std::vector<Node>& top, std::vector<const Node*>& list) {
    if (!text) return SPH_ERR_INVALID_ARGUMENT;
    Reader rd(std::string(text, len));
    top = rd.document();
    list = scene_nodes(top);
    return SPH_OK;
}

// Counter-based uniform in [0, 1) (SplitMix64 finaliser, as sph_scene_lattice).
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline float u01f(uint64_t seed, uint64_t counter) {
    const uint64_t d = mix64(mix64(seed) + (counter + 1) * 0x9E3779B97F4A7C15ull);
    return (float)(d >> 40) * (1.0f / 16777216.0f);
}

}  // namespace

extern "C" {

int sph_xml_scene_count(const char* text, size_t len, uint32_t* count, uint32_t* default_index) {
    if (!count) return SPH_ERR_INVALID_ARGUMENT;
    std::vector<Node> top;
    std::vector<const Node*> list;
    const int st = read_scenes(text, len, top, list);
    if (st != SPH_OK) return st;
    *count = (uint32_t)list.size();
    if (default_index) {
        // the last scene marked default (or def) is the startup scene
        uint32_t d = 0;
        for (size_t k = 0; k < list.size(); ++k)
            if (list[k]->attr("default") || list[k]->attr("def")) d = (uint32_t)k;
        *default_index = d;
    }
    return SPH_OK;
}

int sph_xml_scene(const char* text, size_t len, uint32_t index, sph_scene* scene, sph_scene_info* info) {
    if (!scene) return SPH_ERR_INVALID_ARGUMENT;
    std::vector<Node> top;
    std::vector<const Node*> list;
    const int st = read_scenes(text, len, top, list);
    if (st != SPH_OK) return st;
    SceneVals v;
    defaults(v);
    if (list.empty() && index == 0) {
        // no scene in the file: the default scene (SPH_Scenes.cpp:74-75)
    } else {
        if (index >= list.size()) return SPH_ERR_INVALID_ARGUMENT;
        from_xml(*list[index], v);
    }
    finish(v);
    *scene = v.s;
    if (info) {
        *info = v.info;
        info->chapter = info->chapter || index == 0;  // scenes[0].bChapter = true
    }
    return SPH_OK;
}

int sph_scene_fill(const sph_scene* scene, const sph_scene_info* info, uint64_t seed, float* pos4,
                   float* vel4) {
    if (!scene || !info || !pos4 || !vel4) return SPH_ERR_INVALID_ARGUMENT;
    sph_params p;
    const int st = sph_params_from_scene(scene, &p);
    if (st != SPH_OK) return st;
    const uint32_t n = scene->num_particles;
    const float r = p.particle_r, spc = info->spacing, b = p.dist_bnd_soft;
    if (!(spc > 0.0f)) return SPH_ERR_INVALID_PARAMS;
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 4; ++c) vel4[4 * (size_t)i + c] = 0.0f;
    // worldMinD / worldSizeD of _UpdateGrid, shrunk by r (SPH_Init.cpp:27-31)
    const float bd = b - r;
    float wmin[3], wsize[3];
    for (int c = 0; c < 3; ++c) {
        const float lo = p.world_min[c] + bd, hi = p.world_max[c] - bd;
        wmin[c] = lo + r;
        wsize[c] = (hi - lo) - 2.0f * r;
    }
    if (info->init_type == 1) {
        // random fill of the world; the reference draws rand(), here the
        // library's seeded generator
        for (uint32_t i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c)
                pos4[4 * (size_t)i + c] = wmin[c] + wsize[c] * u01f(seed, 3 * (uint64_t)i + (uint64_t)c);
            pos4[4 * (size_t)i + 3] = 1.0f;
        }
        return SPH_OK;
    }
    // lattice walk from initMin (y raised by r), advancing one axis per site
    // and carrying into the next two in the initLast order (:41-48)
    float imin[3] = {info->init_min[0], info->init_min[1], info->init_min[2]};
    const float imax[3] = {info->init_max[0], info->init_max[1], info->init_max[2]};
    imin[1] = imin[1] + r;
    int ord[3];
    switch (info->init_last) {
        case 0: ord[0] = 1; ord[1] = 2; ord[2] = 0; break;
        case 2: ord[0] = 0; ord[1] = 1; ord[2] = 2; break;
        default: ord[0] = 0; ord[1] = 2; ord[2] = 1; break;
    }
    float x[3] = {imin[0], imin[1], imin[2]};
    auto inc = [&]() -> bool {  // true when the walk wrapped to its start
        const int a = ord[0], bb = ord[1], c = ord[2];
        x[a] = x[a] + spc;
        if (x[a] >= imax[a]) {
            x[a] = imin[a];
            x[bb] = x[bb] + spc;
            if (x[bb] >= imax[bb]) {
                x[bb] = imin[bb];
                x[c] = x[c] + spc;
                if (x[c] >= imax[c]) {
                    x[c] = imin[c];
                    return true;
                }
            }
        }
        return false;
    };
    const bool pump = scene->bnd_type == SPH_BND_PUMP_Y;
    // pump shape (:50-62): rad is the (moved) initMax.x
    const float prad = info->init_max[0], hc = p.h_close - b * 0.90f;
    const float rin = prad * p.pump_rad_in - b / 2.0f;
    const float pxs = (float)std::sin((double)(p.pump_ang_out * p.shape_s[2])) * prad - b / 2.0f - r;
    uint32_t i = 0;
    uint64_t sites_since_hit = 0, period = 0;
    bool counted = false;
    while (i < n) {
        bool keep;
        if (pump) {
            const float ang = p.pump_ang_out;
            keep = (ang < 0.5f && x[1] > 0.0f && x[2] < hc && x[0] < pxs && x[0] > -pxs) ||
                   (ang >= 0.5f && x[1] > 0.0f && x[2] < hc && x[0] > 0.09f * p.shape_s[3]) ||
                   std::sqrt(x[0] * x[0] + x[1] * x[1]) * 1.01f < (x[2] < hc ? prad : rin);
        } else {
            // outside the collider (unless a rotor uses its place), inside the shape (:64-71)
            const float d[3] = {x[0] - p.coll_pos[0], x[1] - p.coll_pos[1], x[2] - p.coll_pos[2]};
            const bool outside = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) > p.coll_r ||
                                 scene->rot_type > 0;
            const uint32_t t = scene->bnd_type;
            keep = outside &&
                   (t == SPH_BND_BOX || info->init_type >= 9 || t == SPH_BND_CYL_YZ ||
                    (t == SPH_BND_CYL_Y && std::sqrt(x[0] * x[0] + x[2] * x[2]) < imax[0]) ||
                    (t == SPH_BND_CYL_Z && std::sqrt(x[0] * x[0] + x[1] * x[1]) < -imin[1]) ||
                    (t == SPH_BND_SPHERE && std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) < -imin[1]));
        }
        if (keep) {
            float* q = pos4 + 4 * (size_t)i;
            q[0] = x[0];
            q[1] = x[1];
            q[2] = x[2];
            q[3] = 1.0f;
            ++i;
            sites_since_hit = 0;
        } else if (++sites_since_hit > period && counted) {
            return SPH_ERR_INVALID_PARAMS;  // a whole pass of the volume kept nothing
        }
        const bool wrapped = inc();
        if (!counted) ++period;
        if (wrapped) counted = true;
    }
    return SPH_OK;
}

}  // extern "C"
