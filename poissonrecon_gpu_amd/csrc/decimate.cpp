// Mesh decimation: EdgeCollapse and TriangleCollapse (reference:
// Geometry.inl:50-254, declared Geometry.cuh:135-137; unused by the
// reference's main()).
//
// Both walk the triangles from the last to the first.  A triangle whose
// corners have already merged is dropped; otherwise, when its sum of squared
// edge lengths exceeds `edge_ratio` times that of an equilateral triangle of
// the same area (12 / sqrt(3) per unit area), its shortest edge (EdgeCollapse)
// or the whole triangle (TriangleCollapse) is merged into the smallest vertex
// index, whose position accumulates the merged positions (averaged at the
// end), and the triangle is removed.  Every decision reads the positions and
// merges of all earlier decisions, and removal swaps the last triangle into
// the removed slot, so the result depends on the exact sequence: the pass is
// sequential by nature and runs on the host, in the reference's float and
// double operations.  Afterwards the surviving vertices are compacted in
// order and every triangle is renumbered (dropping those that became
// degenerate, again by swapping with the last).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/prg.h"

namespace prg {
namespace {

struct V3 {
    float c[3];
};

// SquareDistance / SquareLength (Geometry.inl:29-38): float sums, returned as double
inline double square_distance(const V3& a, const V3& b) {
    const float dx = a.c[0] - b.c[0], dy = a.c[1] - b.c[1], dz = a.c[2] - b.c[2];
    const float s = dx * dx + dy * dy + dz * dz;
    return double(s);
}

inline double length(const V3& p) {
    const float s = p.c[0] * p.c[0] + p.c[1] * p.c[1] + p.c[2] * p.c[2];
    return std::sqrt(double(s));
}

inline int follow(const std::vector<int32_t>& remap, int32_t i) {
    while (remap[i] < i) i = remap[i];
    return i;
}

}  // namespace

// mode 0: EdgeCollapse, 1: TriangleCollapse.  Works on vertices (3 n floats)
// and triangles (3 m ints) in place; returns the new counts.
void decimate(int mode, float edge_ratio, float* vertices, int64_t* n_io, int32_t* triangles, int64_t* m_io) {
    const int64_t n = *n_io;
    std::vector<V3> pos(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) pos[i].c[a] = vertices[3 * i + a];
    std::vector<int32_t> tri(triangles, triangles + 3 * *m_io);
    std::vector<int32_t> remap(static_cast<size_t>(n)), count(static_cast<size_t>(n), 1);
    for (int64_t i = 0; i < n; i++) remap[i] = int32_t(i);
    const double ratio = 12.0 / std::sqrt(3.0);
    auto pop_into = [&](int64_t i) {  // triangles[i] = triangles.back(); pop_back()
        const int64_t last = int64_t(tri.size()) / 3 - 1;
        for (int j = 0; j < 3; j++) tri[3 * i + j] = tri[3 * last + j];
        tri.resize(size_t(3 * last));
    };
    for (int64_t i = int64_t(tri.size()) / 3 - 1; i >= 0; i--) {
        int idx[3];
        for (int j = 0; j < 3; j++) idx[j] = follow(remap, tri[3 * i + j]);
        if (idx[0] == idx[1] || idx[0] == idx[2] || idx[1] == idx[2]) {
            pop_into(i);
            continue;
        }
        V3 p[3], q[2], c;
        double d[3];
        for (int j = 0; j < 3; j++)
            for (int a = 0; a < 3; a++) p[j].c[a] = pos[idx[j]].c[a] / float(count[idx[j]]);
        for (int j = 0; j < 3; j++) {
            q[0].c[j] = p[1].c[j] - p[0].c[j];
            q[1].c[j] = p[2].c[j] - p[0].c[j];
            d[j] = square_distance(p[j], p[(j + 1) % 3]);
        }
        // CrossProduct (Geometry.inl:44-48)
        c.c[0] = q[0].c[1] * q[1].c[2] - q[0].c[2] * q[1].c[1];
        c.c[1] = -q[0].c[0] * q[1].c[2] + q[0].c[2] * q[1].c[0];
        c.c[2] = q[0].c[0] * q[1].c[1] - q[0].c[1] * q[1].c[0];
        const double area = length(c) / 2;
        if (!((d[0] + d[1] + d[2]) * double(edge_ratio) > area * ratio)) continue;
        if (mode == 0) {
            int j = 0;
            if (d[1] < d[j]) j = 1;
            if (d[2] < d[j]) j = 2;
            const int a0 = idx[j], b0 = idx[(j + 1) % 3];
            const int i1 = a0 < b0 ? a0 : b0, i2 = a0 < b0 ? b0 : a0;
            for (int a = 0; a < 3; a++) pos[i1].c[a] += pos[i2].c[a];
            count[i1] += count[i2];
            remap[i2] = i1;
        } else {
            int i1, i2, i3;
            if (idx[0] < idx[1]) {
                if (idx[0] < idx[2]) {
                    i1 = idx[0], i2 = idx[2], i3 = idx[1];
                } else {
                    i1 = idx[2], i2 = idx[0], i3 = idx[1];
                }
            } else {
                if (idx[1] < idx[2]) {
                    i1 = idx[1], i2 = idx[2], i3 = idx[0];
                } else {
                    i1 = idx[2], i2 = idx[1], i3 = idx[0];
                }
            }
            for (int a = 0; a < 3; a++) pos[i1].c[a] += pos[i2].c[a] + pos[i3].c[a];
            count[i1] += count[i2] + count[i3];
            remap[i2] = i1;
            remap[i3] = i1;
        }
        pop_into(i);
    }
    // average, compact the surviving vertices (count becomes the new number)
    int32_t kept = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) pos[i].c[a] /= float(count[i]);
        if (remap[i] == i) {
            pos[kept] = pos[i];
            count[i] = kept;
            kept++;
        }
    }
    for (int64_t i = int64_t(tri.size()) / 3 - 1; i >= 0; i--) {
        int idx[3];
        for (int j = 0; j < 3; j++) {
            idx[j] = follow(remap, tri[3 * i + j]);
            tri[3 * i + j] = count[idx[j]];
        }
        if (idx[0] == idx[1] || idx[0] == idx[2] || idx[1] == idx[2]) pop_into(i);
    }
    for (int32_t i = 0; i < kept; i++)
        for (int a = 0; a < 3; a++) vertices[3 * i + a] = pos[i].c[a];
    for (size_t k = 0; k < tri.size(); k++) triangles[k] = tri[k];
    *n_io = kept;
    *m_io = int64_t(tri.size()) / 3;
}

}  // namespace prg
