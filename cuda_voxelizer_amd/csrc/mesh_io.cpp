// mesh_io.cpp -- mesh file loading (the reference reads meshes with trimesh2,
// main.cpp:174-175: TriMesh::read + need_faces).
//
// Formats: Wavefront OBJ (v / f records; "i/t/n" references; negative,
// i.e. relative, indices; polygons split into a fan around their first
// vertex), OFF, STL (binary or ASCII; three new vertices per facet) and PLY
// (ascii, binary_little_endian, binary_big_endian; vertex x/y/z of any
// numeric type, face lists of any count/index type).  Coordinates are parsed
// with strtof (correctly rounded to binary32, as a "%f" scanf does).
//
// Output: host arrays allocated with malloc -- vertices float[3*nv] and faces
// int32[3*nf] -- released with vox_free_mesh.
#include <cctype>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cuda_voxelizer_amd/vox_capi.h"
#include "vox_host.h"

namespace voxamd {
namespace {

struct Mesh {
    std::vector<float> v;
    std::vector<int32_t> f;
};

bool read_file(const char* path, std::string& data) {
    FILE* fp = fopen(path, "rb");
    if (!fp) return false;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, fp)) > 0) data.append(buf, n);
    fclose(fp);
    return true;
}

std::string lower_ext(const char* path) {
    const char* dot = strrchr(path, '.');
    std::string e = dot ? dot + 1 : "";
    for (char& c : e) c = (char)tolower((unsigned char)c);
    return e;
}

// line cursor over a text buffer
struct Lines {
    const std::string& s;
    size_t pos = 0;
    explicit Lines(const std::string& str) : s(str) {}
    bool next(std::string& line) {
        if (pos >= s.size()) return false;
        size_t e = s.find('\n', pos);
        if (e == std::string::npos) e = s.size();
        line.assign(s, pos, e - pos);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        pos = e + 1;
        return true;
    }
};

void add_polygon(Mesh& m, const std::vector<int32_t>& poly) {
    for (size_t i = 2; i < poly.size(); i++) {
        m.f.push_back(poly[0]);
        m.f.push_back(poly[i - 1]);
        m.f.push_back(poly[i]);
    }
}

std::string load_obj(const std::string& data, Mesh& m) {
    Lines L(data);
    std::string line;
    std::vector<int32_t> poly;
    while (L.next(line)) {
        const char* p = line.c_str();
        while (*p == ' ' || *p == '\t') p++;
        if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
            char* e = (char*)p + 1;
            float xyz[3];
            for (int k = 0; k < 3; k++) {
                const char* b = e;
                xyz[k] = strtof(b, &e);
                if (e == b) return "malformed vertex record: " + line;
            }
            m.v.insert(m.v.end(), xyz, xyz + 3);
        } else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
            poly.clear();
            const char* q = p + 1;
            const long nv = (long)(m.v.size() / 3);
            for (;;) {
                while (*q == ' ' || *q == '\t') q++;
                if (!*q) break;
                char* e;
                const long idx = strtol(q, &e, 10);
                if (e == q) return "malformed face record: " + line;
                const long i0 = idx < 0 ? nv + idx : idx - 1;  // OBJ indices are 1-based, negative = relative
                if (idx == 0 || i0 < 0 || i0 > INT32_MAX) return "face index out of range: " + line;
                poly.push_back((int32_t)i0);
                q = e;
                while (*q && *q != ' ' && *q != '\t') q++;  // skip "/t/n"
            }
            add_polygon(m, poly);
        }
    }
    return "";
}

std::string load_off(const std::string& data, Mesh& m) {
    // whitespace-separated tokens, '#' comments
    std::vector<std::string> tok;
    Lines L(data);
    std::string line;
    while (L.next(line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && isspace((unsigned char)line[i])) i++;
            size_t j = i;
            while (j < line.size() && !isspace((unsigned char)line[j])) j++;
            if (j > i) tok.emplace_back(line, i, j - i);
            i = j;
        }
    }
    size_t t = 0;
    if (tok.empty() || tok[0].find("OFF") == std::string::npos) return "missing OFF header";
    t = 1;
    if (tok.size() < 4) return "truncated OFF header";
    const long nv = strtol(tok[t].c_str(), nullptr, 10), nf = strtol(tok[t + 1].c_str(), nullptr, 10);
    t += 3;
    if (nv < 0 || nf < 0 || (size_t)nv > tok.size() / 3 || tok.size() < t + 3 * (size_t)nv)
        return "truncated OFF vertex list";
    for (long i = 0; i < 3 * nv; i++) m.v.push_back(strtof(tok[t++].c_str(), nullptr));
    std::vector<int32_t> poly;
    for (long i = 0; i < nf; i++) {
        if (t >= tok.size()) return "truncated OFF face list";
        const long k = strtol(tok[t++].c_str(), nullptr, 10);
        if (k < 0 || t + (size_t)k > tok.size()) return "truncated OFF face list";
        poly.clear();
        for (long j = 0; j < k; j++) {
            const long idx = strtol(tok[t++].c_str(), nullptr, 10);
            if (idx < 0 || idx >= nv) return "OFF face index out of range";
            poly.push_back((int32_t)idx);
        }
        add_polygon(m, poly);
    }
    return "";
}

std::string load_stl(const std::string& data, Mesh& m) {
    if (data.size() >= 84) {
        uint32_t n;
        memcpy(&n, data.data() + 80, 4);
        if (84ull + 50ull * n == data.size()) {  // binary
            for (uint32_t i = 0; i < n; i++) {
                const char* rec = data.data() + 84 + 50ull * i;
                for (int k = 0; k < 3; k++) {
                    float xyz[3];
                    memcpy(xyz, rec + 12 + 12 * k, 12);
                    m.v.insert(m.v.end(), xyz, xyz + 3);
                    m.f.push_back((int32_t)(m.v.size() / 3 - 1));
                }
            }
            return "";
        }
    }
    Lines L(data);
    std::string line;
    while (L.next(line)) {
        const char* p = line.c_str();
        while (*p == ' ' || *p == '\t') p++;
        if (strncmp(p, "vertex", 6) == 0) {
            char* e = (char*)p + 6;
            for (int k = 0; k < 3; k++) {
                const char* b = e;
                m.v.push_back(strtof(b, &e));
                if (e == b) return "malformed STL vertex: " + line;
            }
        }
    }
    if (m.v.size() % 9) return "STL vertex count is not a multiple of 3";
    for (size_t i = 0; i < m.v.size() / 3; i++) m.f.push_back((int32_t)i);
    return "";
}

// ---- PLY ----
enum PType { P_NONE, P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64 };

PType ptype(const std::string& s) {
    if (s == "char" || s == "int8") return P_I8;
    if (s == "uchar" || s == "uint8") return P_U8;
    if (s == "short" || s == "int16") return P_I16;
    if (s == "ushort" || s == "uint16") return P_U16;
    if (s == "int" || s == "int32") return P_I32;
    if (s == "uint" || s == "uint32") return P_U32;
    if (s == "float" || s == "float32") return P_F32;
    if (s == "double" || s == "float64") return P_F64;
    return P_NONE;
}

int psize(PType t) {
    switch (t) {
        case P_I8: case P_U8: return 1;
        case P_I16: case P_U16: return 2;
        case P_I32: case P_U32: case P_F32: return 4;
        case P_F64: return 8;
        default: return 0;
    }
}

struct PProp {
    std::string name;
    PType type = P_NONE, count = P_NONE;  // count != P_NONE: a list
};
struct PElem {
    std::string name;
    long n = 0;
    std::vector<PProp> props;
};

// binary64 -> binary32 rounded to nearest, without a conversion whose source
// lies outside the float range (undefined in C++): values from FLT_MAX + half
// an ulp up round to infinity, those just above FLT_MAX to FLT_MAX.
float to_float(double d) {
    const double top = (double)FLT_MAX, lim = top + 0x1p103;
    if (d >= lim) return INFINITY;
    if (d <= -lim) return -INFINITY;
    if (d > top) return FLT_MAX;
    if (d < -top) return -FLT_MAX;
    return (float)d;  // NaN passes through
}

// binary reader: returns the value as double (exact for every PLY type)
struct Bin {
    const std::string& d;
    size_t pos;
    bool big;
    bool ok = true;
    double get(PType t) {
        const int n = psize(t);
        if (pos + n > d.size()) {
            ok = false;
            return 0;
        }
        unsigned char b[8];
        memcpy(b, d.data() + pos, n);
        pos += n;
        if (big)
            for (int i = 0; i < n / 2; i++) std::swap(b[i], b[n - 1 - i]);
        switch (t) {
            case P_I8: { int8_t v; memcpy(&v, b, 1); return v; }
            case P_U8: return b[0];
            case P_I16: { int16_t v; memcpy(&v, b, 2); return v; }
            case P_U16: { uint16_t v; memcpy(&v, b, 2); return v; }
            case P_I32: { int32_t v; memcpy(&v, b, 4); return v; }
            case P_U32: { uint32_t v; memcpy(&v, b, 4); return v; }
            case P_F32: { float v; memcpy(&v, b, 4); return v; }
            case P_F64: { double v; memcpy(&v, b, 8); return v; }
            default: ok = false; return 0;
        }
    }
};

std::string load_ply(const std::string& data, Mesh& m) {
    Lines L(data);
    std::string line;
    if (!L.next(line) || line.compare(0, 3, "ply") != 0) return "missing ply magic";
    int fmt = -1;  // 0 ascii, 1 little, 2 big
    std::vector<PElem> elems;
    while (L.next(line)) {
        char a[64] = {0}, b[64] = {0}, c[64] = {0}, d[64] = {0}, e[64] = {0};
        const int k = sscanf(line.c_str(), "%63s %63s %63s %63s %63s", a, b, c, d, e);
        if (k <= 0) continue;
        if (!strcmp(a, "format")) {
            fmt = !strcmp(b, "ascii") ? 0 : !strcmp(b, "binary_little_endian") ? 1 : !strcmp(b, "binary_big_endian") ? 2 : -1;
        } else if (!strcmp(a, "element") && k >= 3) {
            PElem el;
            el.name = b;
            el.n = strtol(c, nullptr, 10);
            elems.push_back(el);
        } else if (!strcmp(a, "property") && !elems.empty()) {
            PProp p;
            if (!strcmp(b, "list") && k >= 5) {
                p.count = ptype(c);
                p.type = ptype(d);
                p.name = e;
            } else if (k >= 3) {
                p.type = ptype(b);
                p.name = c;
            }
            if (p.type == P_NONE) return "unsupported PLY property type: " + line;
            elems.back().props.push_back(p);
        } else if (!strcmp(a, "end_header")) {
            break;
        }
    }
    if (fmt < 0) return "unsupported PLY format";
    std::vector<int32_t> poly;
    Bin bin{data, L.pos, fmt == 2};
    Lines A(data);
    A.pos = L.pos;
    std::vector<std::string> toks;
    size_t ti = 0;
    // next ascii token (false at the end of the data)
    auto next_tok = [&](std::string& t) -> bool {
        while (ti >= toks.size()) {
            std::string l;
            if (!A.next(l)) return false;
            toks.clear();
            ti = 0;
            size_t i = 0;
            while (i < l.size()) {
                while (i < l.size() && isspace((unsigned char)l[i])) i++;
                size_t j = i;
                while (j < l.size() && !isspace((unsigned char)l[j])) j++;
                if (j > i) toks.emplace_back(l, i, j - i);
                i = j;
            }
        }
        t = toks[ti++];
        return true;
    };
    std::string tok;
    for (const PElem& el : elems) {
        const bool is_v = el.name == "vertex", is_f = el.name == "face";
        int ix = -1, iy = -1, iz = -1;
        for (size_t i = 0; i < el.props.size(); i++) {
            if (el.props[i].name == "x") ix = (int)i;
            if (el.props[i].name == "y") iy = (int)i;
            if (el.props[i].name == "z") iz = (int)i;
        }
        if (is_v && (ix < 0 || iy < 0 || iz < 0)) return "PLY vertex element without x/y/z";
        if (el.props.empty()) continue;  // rows of nothing: no data to skip
        for (long r = 0; r < el.n; r++) {
            float xyz[3] = {0, 0, 0};
            for (size_t i = 0; i < el.props.size(); i++) {
                const PProp& p = el.props[i];
                if (p.count != P_NONE) {
                    double cnt;
                    if (fmt) {
                        cnt = bin.get(p.count);
                    } else {
                        if (!next_tok(tok)) return "truncated PLY body";
                        cnt = strtod(tok.c_str(), nullptr);
                    }
                    if (!bin.ok || !(cnt >= 0) || cnt > 1e6) return "bad PLY list";
                    poly.clear();
                    for (long j = 0; j < (long)cnt; j++) {
                        double idx;
                        if (fmt) {
                            idx = bin.get(p.type);
                        } else {
                            if (!next_tok(tok)) return "truncated PLY body";
                            idx = strtod(tok.c_str(), nullptr);
                        }
                        if (!bin.ok) return "truncated PLY body";
                        if (!(idx >= 0.0 && idx < 2147483648.0)) return "PLY face index out of range";
                        poly.push_back((int32_t)idx);
                    }
                    if (!bin.ok) return "truncated PLY body";
                    if (is_f && (p.name == "vertex_indices" || p.name == "vertex_index")) add_polygon(m, poly);
                } else {
                    float v;
                    if (fmt) {
                        v = to_float(bin.get(p.type));  // binary32 stays exact; binary64 rounds once
                        if (!bin.ok) return "truncated PLY body";
                    } else {
                        if (!next_tok(tok)) return "truncated PLY body";
                        v = strtof(tok.c_str(), nullptr);
                    }
                    if ((int)i == ix) xyz[0] = v;
                    if ((int)i == iy) xyz[1] = v;
                    if ((int)i == iz) xyz[2] = v;
                }
            }
            if (is_v) m.v.insert(m.v.end(), xyz, xyz + 3);
        }
    }
    const long nv = (long)(m.v.size() / 3);
    for (int32_t idx : m.f)
        if (idx < 0 || idx >= nv) return "PLY face index out of range";
    return "";
}

}  // namespace
}  // namespace voxamd

using namespace voxamd;

extern "C" {

vox_status vox_load_mesh(const char* path, vox_mesh* out) {
    if (!path || !out) return fail(VOX_ERR_INVALID, "NULL argument");
    memset(out, 0, sizeof *out);
    std::string data;
    if (!read_file(path, data)) return fail(VOX_ERR_INVALID, std::string("cannot read mesh file: ") + path);
    Mesh m;
    const std::string ext = lower_ext(path);
    std::string err;
    if (ext == "obj")
        err = load_obj(data, m);
    else if (ext == "off")
        err = load_off(data, m);
    else if (ext == "stl")
        err = load_stl(data, m);
    else if (ext == "ply")
        err = load_ply(data, m);
    else
        return fail(VOX_ERR_UNSUPPORTED, "unsupported mesh format ." + ext + " (obj, off, stl, ply)");
    if (!err.empty()) return fail(VOX_ERR_INVALID, std::string(path) + ": " + err);
    const long nv = (long)(m.v.size() / 3);
    for (int32_t idx : m.f)
        if (idx < 0 || idx >= nv) return fail(VOX_ERR_INVALID, std::string(path) + ": face index out of range");
    out->n_vertices = m.v.size() / 3;
    out->n_faces = m.f.size() / 3;
    out->vertices = (float*)malloc(m.v.size() * sizeof(float) + 4);
    out->faces = (int32_t*)malloc(m.f.size() * sizeof(int32_t) + 4);
    if (!out->vertices || !out->faces) {
        free(out->vertices);
        free(out->faces);
        memset(out, 0, sizeof *out);
        return fail(VOX_ERR_OOM, "host allocation for the mesh");
    }
    if (!m.v.empty()) memcpy(out->vertices, m.v.data(), m.v.size() * sizeof(float));
    if (!m.f.empty()) memcpy(out->faces, m.f.data(), m.f.size() * sizeof(int32_t));
    return VOX_OK;
}

void vox_free_mesh(vox_mesh* mesh) {
    if (!mesh) return;
    free(mesh->vertices);
    free(mesh->faces);
    memset(mesh, 0, sizeof *mesh);
}

}  // extern "C"
