// Scene ingest: Wavefront OBJ files to the triangle soup the builders take,
// with the semantics of the reference framework's importer
// (gpu-ray-traversal/src/framework/io/MeshWavefrontIO.cpp: loadObj :330-520,
// importWavefrontMesh :527-545), of Scene's triangle list
// (gpu-ray-traversal/src/rt/Scene.cpp:55-77) and of
// SceneConverter::ConvertToCudaSceneDeviceMemory
// (gpu-ray-traversal/src/rt/SceneConverter.cpp:19-71).  Host code only.
//
// Kept from the reference because they decide which triangles a scene has,
// in which order, and their exact coordinates:
//   - lines are read with trailing whitespace trimmed and '\' continuations
//     joined; a directive is recognised by its literal prefix ("v ", "f ",
//     "usemtl ", "mtllib ") after leading blanks;
//   - "v" needs exactly three numbers, otherwise the line is ignored (and no
//     position is added);
//   - numbers are parsed by the framework's own accumulation in fp32
//     (String.cpp:440-491: digits * 10 + d, fraction digits scaled by
//     repeated * 0.1f, exponent * powf(10, e)), not by a correctly rounded
//     conversion;
//   - face corners "p", "p/t", "p/t/n", "p//n"; negative indices count back
//     from the current end, out-of-range positions give (0, 0, 0); polygons
//     are fans (v0, v[i-1], v[i]); a face line that does not parse to its end
//     adds nothing;
//   - faces are grouped by material: "usemtl" of a material defined in a
//     loaded "mtllib" file sends faces to that material's group (created at
//     its first use), any other "usemtl" and faces before the first one go to
//     a default group created at the first such face; the scene lists the
//     groups in creation order, each in file order;
//   - the converter's scene box starts its maximum at FLT_MIN.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

struct atr_mesh {
    std::vector<float> positions;              // 3 per position
    std::vector<std::vector<int32_t>> groups;  // 3 position indices (or -1) per triangle
};

namespace atr {

namespace {

// The whole file, read to its end (no size taken from seeking, which a
// directory or a pipe does not give); false if it cannot be opened or read.
bool read_file(const std::string& path, std::string& out)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    char buf[1 << 16];
    for (;;) {
        const size_t got = std::fread(buf, 1, sizeof buf, f);
        out.append(buf, got);
        if (got < sizeof buf) break;  // end of file or a read error
    }
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

// Lines with '\r' dropped, '\' + newline joining the next line, trailing
// blanks trimmed (BufferedInputStream::readLine(true, true)).
std::vector<std::string> split_lines(const std::string& text)
{
    std::vector<std::string> lines;
    std::string cur;
    for (size_t i = 0; i < text.size(); ++i) {
        const char c = text[i];
        if (c == '\r') continue;
        if (c == '\n') {
            if (!cur.empty() && cur.back() == '\\') {
                cur.pop_back();
                continue;
            }
            lines.push_back(cur);
            cur.clear();
            continue;
        }
        cur.push_back(c);
    }
    if (!cur.empty()) lines.push_back(cur);
    for (std::string& l : lines) {
        size_t e = l.size();
        while (e > 0 && (l[e - 1] == ' ' || l[e - 1] == '\t')) --e;
        l.resize(e);
    }
    return lines;
}

void parse_space(const char*& p)
{
    while (*p == ' ' || *p == '\t') ++p;
}

bool parse_literal(const char*& p, const char* lit)
{
    const char* t = p;
    while (*lit && *t == *lit) {
        ++t;
        ++lit;
    }
    if (*lit) return false;
    p = t;
    return true;
}

bool parse_char(const char*& p, char c)
{
    if (*p != c) return false;
    ++p;
    return true;
}

// Decimal integer (String.cpp:384-399).  The digits accumulate modulo 2^32,
// the two's-complement wrap of the framework's int32 accumulation, without
// signed overflow: a run of digits too long for int32 gives the same bits.
bool parse_int(const char*& p, int32_t& value)
{
    const char* t = p;
    uint32_t v = 0;
    const bool neg = (!parse_char(t, '+') && parse_char(t, '-'));
    if (*t < '0' || *t > '9') return false;
    while (*t >= '0' && *t <= '9') v = v * 10u + (uint32_t)(*t++ - '0');
    value = (int32_t)(neg ? 0u - v : v);
    p = t;
    return true;
}

// The framework's fp32 number parser (String.cpp:440-491).
bool parse_float(const char*& p, float& value)
{
    const char* t = p;
    const bool neg = (!parse_char(t, '+') && parse_char(t, '-'));
    float v = 0.0f;
    int digits = 0;
    while (*t >= '0' && *t <= '9') {
        v = v * 10.0f + (float)(*t++ - '0');
        ++digits;
    }
    if (parse_char(t, '.')) {
        float scale = 1.0f;
        while (*t >= '0' && *t <= '9') {
            scale *= 0.1f;
            v += scale * (float)(*t++ - '0');
            ++digits;
        }
    }
    if (!digits) return false;
    p = t;
    if (*p == '#') {  // the framework's spellings of non-finite values
        const char* q = p;
        if (parse_literal(q, "#INF")) {
            value = neg ? -INFINITY : INFINITY;
            p = q;
            return true;
        }
        if (parse_literal(q, "#SNAN") || parse_literal(q, "#QNAN") || parse_literal(q, "#IND")) {
            value = neg ? -NAN : NAN;
            p = q;
            return true;
        }
    }
    int32_t e = 0;
    if ((parse_char(t, 'e') || parse_char(t, 'E')) && parse_int(t, e)) {
        p = t;
        if (e) v *= std::pow(10.0f, (float)e);
    }
    value = neg ? -v : v;
    return true;
}

bool parse_floats(const char*& p, float* out, int count)
{
    const char* t = p;
    for (int i = 0; i < count; ++i) {
        parse_space(t);
        if (!parse_float(t, out[i])) return false;
    }
    p = t;
    return true;
}

std::string dir_of(const std::string& path)
{
    const size_t k = path.find_last_of("/\\");
    return k == std::string::npos ? std::string() : path.substr(0, k);
}

// Names defined by "newmtl" in a material library (the other statements
// only set shading parameters and do not change the geometry).
void load_mtl_names(const std::string& path, std::map<std::string, int>& materials)
{
    std::string text;
    if (!read_file(path, text)) return;
    for (const std::string& line : split_lines(text)) {
        const char* p = line.c_str();
        parse_space(p);
        if (parse_literal(p, "newmtl ")) {
            parse_space(p);
            if (*p && !materials.count(p)) materials.emplace(p, -1);
        }
    }
}

}  // namespace

atr_status load_obj(const std::string& path, atr_mesh* m)
{
    std::string text;
    if (!read_file(path, text)) return fail(ATR_ERR_INVALID_ARGUMENT, "cannot read " + path);
    const std::string dir = dir_of(path);
    std::map<std::string, int> materials;  // name -> group (-1 until first use)
    int group = -1, default_group = -1;
    size_t texcoords = 0, normals = 0;
    std::vector<int32_t> corners;
    for (const std::string& line : split_lines(text)) {
        const char* p = line.c_str();
        parse_space(p);
        if (!*p || parse_literal(p, "#")) continue;
        if (parse_literal(p, "v ")) {
            parse_space(p);
            float v[3];
            if (parse_floats(p, v, 3)) {
                parse_space(p);
                if (!*p) m->positions.insert(m->positions.end(), v, v + 3);
            }
        } else if (parse_literal(p, "vt ")) {
            parse_space(p);
            float v[2], dummy;
            if (parse_floats(p, v, 2)) {
                parse_space(p);
                while (parse_float(p, dummy)) parse_space(p);
                if (!*p) ++texcoords;
            }
        } else if (parse_literal(p, "vn ")) {
            parse_space(p);
            float v[3];
            if (parse_floats(p, v, 3)) {
                parse_space(p);
                if (!*p) ++normals;
            }
        } else if (parse_literal(p, "f ")) {
            parse_space(p);
            corners.clear();
            while (*p) {
                int32_t ptn[3] = {0, 0, 0};
                if (!parse_int(p, ptn[0])) break;
                for (int i = 1; i < 4 && parse_literal(p, "/"); ++i) {
                    int32_t tmp = 0;
                    parse_int(p, tmp);
                    if (i < 3) ptn[i] = tmp;
                }
                parse_space(p);
                const int64_t size[3] = {(int64_t)(m->positions.size() / 3), (int64_t)texcoords,
                                         (int64_t)normals};
                int64_t pos = ptn[0];
                pos = pos < 0 ? pos + size[0] : pos - 1;
                if (pos < 0 || pos >= size[0]) pos = -1;
                corners.push_back((int32_t)pos);
            }
            if (!*p) {
                if (group == -1) {
                    if (default_group == -1) {
                        default_group = (int)m->groups.size();
                        m->groups.emplace_back();
                    }
                    group = default_group;
                }
                std::vector<int32_t>& g = m->groups[group];
                for (size_t i = 2; i < corners.size(); ++i) {
                    g.push_back(corners[0]);
                    g.push_back(corners[i - 1]);
                    g.push_back(corners[i]);
                }
            }
        } else if (parse_literal(p, "usemtl ")) {
            parse_space(p);
            group = -1;
            auto it = materials.find(p);
            if (it != materials.end()) {
                if (it->second == -1) {
                    it->second = (int)m->groups.size();
                    m->groups.emplace_back();
                }
                group = it->second;
            }
        } else if (parse_literal(p, "mtllib ")) {
            parse_space(p);
            // the reference reads material libraries only for a file named
            // with a directory (MeshWavefrontIO.cpp:462)
            if (*p && !dir.empty()) load_mtl_names(dir + "/" + p, materials);
        }
    }
    return ATR_OK;
}

}  // namespace atr

using namespace atr;

extern "C" {

atr_status atr_mesh_load_obj(const char* path, atr_mesh** out)
try {
    if (!path || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    std::unique_ptr<atr_mesh> m(new atr_mesh());
    ATR_TRY(load_obj(path, m.get()));
    *out = m.release();
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_mesh_info(const atr_mesh* m, uint32_t* num_triangles, uint32_t* num_positions)
try {
    if (!m) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    uint64_t tris = 0;
    for (const auto& g : m->groups) tris += g.size() / 3;
    if (tris >= (1ull << 30)) return fail(ATR_ERR_UNSUPPORTED, "too many triangles");
    if (num_triangles) *num_triangles = (uint32_t)tris;
    if (num_positions) *num_positions = (uint32_t)(m->positions.size() / 3);
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_mesh_soup(const atr_mesh* m, float* vertices)
try {
    if (!m || !vertices) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    size_t k = 0;
    for (const auto& g : m->groups) {
        for (const int32_t idx : g) {
            float* q = vertices + 4 * k++;
            if (idx < 0) {
                q[0] = q[1] = q[2] = 0.0f;
            } else {
                q[0] = m->positions[3ull * idx];
                q[1] = m->positions[3ull * idx + 1];
                q[2] = m->positions[3ull * idx + 2];
            }
            q[3] = 0.0f;
        }
    }
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_mesh_destroy(atr_mesh* m)
try {
    delete m;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_scene_create_from_mesh(atr_context* ctx, const atr_mesh* m, atr_scene** out)
try {
    if (!ctx || !m || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    uint32_t n = 0;
    ATR_TRY(atr_mesh_info(m, &n, nullptr));
    if (n < 2) return fail(ATR_ERR_INVALID_ARGUMENT, "the mesh has fewer than 2 triangles");
    std::vector<float> v(12ull * n);
    ATR_TRY(atr_mesh_soup(m, v.data()));
    // SceneConverter.cpp:28-63: running min from FLT_MAX, max from FLT_MIN
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {FLT_MIN, FLT_MIN, FLT_MIN};
    for (size_t i = 0; i < 3ull * n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float x = v[4 * i + a];
            if (x < mn[a]) mn[a] = x;
            if (x > mx[a]) mx[a] = x;
        }
    return atr_scene_create(ctx, v.data(), n, mn, mx, out);
}
ATR_CATCH_ALL

}  // extern "C"
