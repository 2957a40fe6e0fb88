// Point-cloud readers and the ASCII PLY mesh writer (reference:
// PointStream.inl:29-239 readers, plyfile.cu:2769-2837 PlyWriteTriangles).
//   .ply   vertex element with float properties x y z nx ny nz (any order,
//          other properties skipped); ascii, binary_little_endian or
//          binary_big_endian;
//   .bnpts raw little-endian float32 x y z nx ny nz per point;
//   other  ASCII, six floats per point.
// The writer emits "ply / format ascii 1.0", vertex x y z as "%g" of
// v * scale + center and faces as "3 i j k".
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prg.h"

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(const char* p, const char* m) : f(std::fopen(p, m)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

std::string ext_of(const char* path) {
    const char* dot = std::strrchr(path, '.');
    std::string e = dot ? std::string(dot + 1) : std::string();
    for (auto& ch : e) ch = char(std::tolower(ch));
    return e;
}

int type_size(const std::string& t) {
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
    if (t == "double" || t == "float64") return 8;
    return 0;
}

double read_bin(const unsigned char* p, const std::string& t, bool swap) {
    unsigned char b[8];
    const int n = type_size(t);
    for (int i = 0; i < n; i++) b[i] = swap ? p[n - 1 - i] : p[i];
    if (t == "float" || t == "float32") {
        float v;
        std::memcpy(&v, b, 4);
        return v;
    }
    if (t == "double" || t == "float64") {
        double v;
        std::memcpy(&v, b, 8);
        return v;
    }
    if (t == "int" || t == "int32") {
        int32_t v;
        std::memcpy(&v, b, 4);
        return v;
    }
    if (t == "uint" || t == "uint32") {
        uint32_t v;
        std::memcpy(&v, b, 4);
        return v;
    }
    if (t == "short" || t == "int16") {
        int16_t v;
        std::memcpy(&v, b, 2);
        return v;
    }
    if (t == "ushort" || t == "uint16") {
        uint16_t v;
        std::memcpy(&v, b, 2);
        return v;
    }
    if (t == "char" || t == "int8") return double(int8_t(b[0]));
    return double(b[0]);
}

prg_status read_ply(FILE* f, std::vector<float>& P, std::vector<float>& Nn) {
    char line[1024];
    if (!std::fgets(line, sizeof(line), f) || std::strncmp(line, "ply", 3) != 0) return PRG_E_IO;
    std::string format;
    int64_t nvert = -1;
    bool in_vertex = false, seen_vertex = false;
    struct Prop {
        std::string name, type;
    };
    std::vector<Prop> props;
    std::vector<std::pair<std::string, int64_t>> before;  // elements preceding "vertex"
    while (std::fgets(line, sizeof(line), f)) {
        char a[256] = {0}, b[256] = {0}, c[256] = {0};
        int k = std::sscanf(line, "%255s %255s %255s", a, b, c);
        if (k <= 0) continue;
        std::string w = a;
        if (w == "format") {
            format = b;
        } else if (w == "element") {
            in_vertex = (std::string(b) == "vertex");
            if (in_vertex) {
                nvert = std::atoll(c);
                seen_vertex = true;
            } else if (!seen_vertex) {
                before.emplace_back(b, std::atoll(c));
            }
        } else if (w == "property" && in_vertex) {
            if (std::string(b) == "list") return PRG_E_IO;
            props.push_back({c, b});
        } else if (w == "end_header") {
            break;
        }
    }
    if (nvert < 1 || !before.empty()) return PRG_E_IO;
    int ix[6] = {-1, -1, -1, -1, -1, -1};
    const char* names[6] = {"x", "y", "z", "nx", "ny", "nz"};
    for (size_t i = 0; i < props.size(); i++)
        for (int k = 0; k < 6; k++)
            if (props[i].name == names[k]) ix[k] = int(i);
    for (int k = 0; k < 6; k++)
        if (ix[k] < 0) return PRG_E_IO;  // x y z nx ny nz required (PointStream.inl:112-120)
    P.resize(size_t(3 * nvert));
    Nn.resize(size_t(3 * nvert));
    if (format == "ascii") {
        std::vector<double> v(props.size());
        for (int64_t i = 0; i < nvert; i++) {
            for (size_t k = 0; k < props.size(); k++)
                if (std::fscanf(f, "%lf", &v[k]) != 1) return PRG_E_IO;
            for (int a = 0; a < 3; a++) {
                P[3 * i + a] = float(v[ix[a]]);
                Nn[3 * i + a] = float(v[ix[3 + a]]);
            }
        }
        return PRG_OK;
    }
    const bool be = (format == "binary_big_endian");
    if (!be && format != "binary_little_endian") return PRG_E_IO;
    std::vector<int> off(props.size());
    int stride = 0;
    for (size_t k = 0; k < props.size(); k++) {
        const int s = type_size(props[k].type);
        if (!s) return PRG_E_IO;
        off[k] = stride;
        stride += s;
    }
    std::vector<unsigned char> rec(static_cast<size_t>(stride));
    for (int64_t i = 0; i < nvert; i++) {
        if (std::fread(rec.data(), 1, size_t(stride), f) != size_t(stride)) return PRG_E_IO;
        for (int a = 0; a < 3; a++) {
            P[3 * i + a] = float(read_bin(&rec[off[ix[a]]], props[ix[a]].type, be));
            Nn[3 * i + a] = float(read_bin(&rec[off[ix[3 + a]]], props[ix[3 + a]].type, be));
        }
    }
    return PRG_OK;
}

}  // namespace

extern "C" {

prg_status prg_read_points(const char* path, float** xyz, float** nxyz, int64_t* n) {
    if (!path || !xyz || !nxyz || !n) return PRG_E_ARG;
    *xyz = *nxyz = nullptr;
    *n = 0;
    const std::string e = ext_of(path);
    std::vector<float> P, Nn;
    if (e == "ply") {
        File f(path, "rb");
        if (!f.f) return PRG_E_IO;
        prg_status st = read_ply(f.f, P, Nn);
        if (st != PRG_OK) return st;
    } else if (e == "bnpts") {
        File f(path, "rb");
        if (!f.f) return PRG_E_IO;
        float rec[6];
        size_t got;
        while ((got = std::fread(rec, sizeof(float), 6, f.f)) == 6)
            for (int a = 0; a < 3; a++) {
                P.push_back(rec[a]);
                Nn.push_back(rec[3 + a]);
            }
        if (got != 0) return PRG_E_IO;  // truncated record
    } else {
        File f(path, "r");
        if (!f.f) return PRG_E_IO;
        double v[6];
        int k;
        while ((k = std::fscanf(f.f, "%lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5])) == 6)
            for (int a = 0; a < 3; a++) {
                P.push_back(float(v[a]));
                Nn.push_back(float(v[3 + a]));
            }
        if (k != EOF) return PRG_E_IO;
    }
    const int64_t cnt = int64_t(P.size() / 3);
    if (cnt < 1) return PRG_E_IO;
    *xyz = static_cast<float*>(std::malloc(P.size() * 4));
    *nxyz = static_cast<float*>(std::malloc(Nn.size() * 4));
    if (!*xyz || !*nxyz) {
        std::free(*xyz);
        std::free(*nxyz);
        *xyz = *nxyz = nullptr;
        return PRG_E_NOMEM;
    }
    std::memcpy(*xyz, P.data(), P.size() * 4);
    std::memcpy(*nxyz, Nn.data(), Nn.size() * 4);
    *n = cnt;
    return PRG_OK;
}

void prg_points_release(float* xyz, float* nxyz) {
    std::free(xyz);
    std::free(nxyz);
}

prg_status prg_write_mesh_ply(const char* path, const prg_mesh* m) {
    if (!path || !m) return PRG_E_ARG;
    File f(path, "w");
    if (!f.f) return PRG_E_IO;
    std::fprintf(f.f, "ply\nformat ascii 1.0\nelement vertex %lld\nproperty float x\nproperty float y\n"
                      "property float z\nelement face %lld\nproperty list uchar int vertex_indices\nend_header\n",
                 (long long)m->n_vertices, (long long)m->n_triangles);
    for (int64_t i = 0; i < m->n_vertices; i++) {
        float o[3];
        for (int a = 0; a < 3; a++) {
            const float t = m->vertices[3 * i + a] * m->scale;
            o[a] = t + m->center[a];
        }
        // every item is followed by one space, every element by a newline
        // (ply_put_element / write_ascii_item, plyfile.cu:597-626,2117-2147)
        std::fprintf(f.f, "%g %g %g \n", double(o[0]), double(o[1]), double(o[2]));
    }
    for (int64_t i = 0; i < m->n_triangles; i++)
        std::fprintf(f.f, "3 %d %d %d \n", m->triangles[3 * i], m->triangles[3 * i + 1], m->triangles[3 * i + 2]);
    return std::ferror(f.f) ? PRG_E_IO : PRG_OK;
}

}  // extern "C"
