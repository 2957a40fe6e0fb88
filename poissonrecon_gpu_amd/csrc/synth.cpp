// Seeded synthetic workloads (include/prg_synth.h).  Host code, double
// precision internally, float outputs.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/prg_synth.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
    uint64_t next() {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {  // Box-Muller, one value per call
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

void unit(double v[3]) {
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0)
        for (int a = 0; a < 3; a++) v[a] /= l;
}

// Tilt unit normal n by a Gaussian angle of deviation `sigma` about a random tangent axis.
void tilt(Rng& r, double n[3], double sigma) {
    double t[3] = {r.normal(), r.normal(), r.normal()};
    double d = t[0] * n[0] + t[1] * n[1] + t[2] * n[2];
    for (int a = 0; a < 3; a++) t[a] -= d * n[a];
    unit(t);
    const double ang = sigma * r.normal();
    for (int a = 0; a < 3; a++) n[a] = std::cos(ang) * n[a] + std::sin(ang) * t[a];
    unit(n);
}

struct SQ {  // superquadric: |x/a|^(2/e2)+|y/b|^(2/e2))^(e2/e1) + |z/c|^(2/e1) = 1
    double e1, e2, ax[3], c[3], R[3][3];  // R rows: local axes in world
    double area = 0.0;
    // inside-outside function in local coordinates
    double F(const double q[3]) const {
        const double x = std::fabs(q[0] / ax[0]), y = std::fabs(q[1] / ax[1]), z = std::fabs(q[2] / ax[2]);
        const double xy = std::pow(std::pow(x, 2.0 / e2) + std::pow(y, 2.0 / e2), e2 / e1);
        return xy + std::pow(z, 2.0 / e1);
    }
    void to_local(const double p[3], double q[3]) const {
        double d[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
        for (int i = 0; i < 3; i++) q[i] = R[i][0] * d[0] + R[i][1] * d[1] + R[i][2] * d[2];
    }
    bool inside_world(const double p[3]) const {
        double q[3];
        to_local(p, q);
        return F(q) < 1.0;
    }
    // point on the surface along the ray from the centre with local direction v
    void project(const double v[3], double q[3]) const {
        const double f = F(v);
        const double t = std::pow(f, -e1 / 2.0);
        for (int a = 0; a < 3; a++) q[a] = v[a] * t;
    }
    // outward normal (gradient of F) in local coordinates
    void grad(const double q[3], double g[3]) const {
        const double x = q[0] / ax[0], y = q[1] / ax[1], z = q[2] / ax[2];
        const double ex = std::pow(std::fabs(x), 2.0 / e2), ey = std::pow(std::fabs(y), 2.0 / e2);
        const double base = ex + ey;
        const double pre = (base > 0) ? std::pow(base, e2 / e1 - 1.0) * (2.0 / e1) : 0.0;
        auto sg = [](double v) { return v < 0 ? -1.0 : 1.0; };
        g[0] = (x != 0) ? pre * std::pow(std::fabs(x), 2.0 / e2 - 1.0) * sg(x) / ax[0] : 0.0;
        g[1] = (y != 0) ? pre * std::pow(std::fabs(y), 2.0 / e2 - 1.0) * sg(y) / ax[1] : 0.0;
        g[2] = (z != 0) ? (2.0 / e1) * std::pow(std::fabs(z), 2.0 / e1 - 1.0) * sg(z) / ax[2] : 0.0;
    }
    // parametric point (eta in [-pi/2,pi/2], om in [-pi,pi))
    void param(double eta, double om, double q[3]) const {
        auto spow = [](double v, double e) { return (v < 0 ? -1.0 : 1.0) * std::pow(std::fabs(v), e); };
        const double ce = spow(std::cos(eta), e1), se = spow(std::sin(eta), e1);
        q[0] = ax[0] * ce * spow(std::cos(om), e2);
        q[1] = ax[1] * ce * spow(std::sin(om), e2);
        q[2] = ax[2] * se;
    }
};

void random_rotation(Rng& r, double R[3][3]) {
    // uniform random quaternion
    double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    double qw = std::sqrt(1 - u1) * std::sin(2 * M_PI * u2), qx = std::sqrt(1 - u1) * std::cos(2 * M_PI * u2);
    double qy = std::sqrt(u1) * std::sin(2 * M_PI * u3), qz = std::sqrt(u1) * std::cos(2 * M_PI * u3);
    R[0][0] = 1 - 2 * (qy * qy + qz * qz);
    R[0][1] = 2 * (qx * qy - qz * qw);
    R[0][2] = 2 * (qx * qz + qy * qw);
    R[1][0] = 2 * (qx * qy + qz * qw);
    R[1][1] = 1 - 2 * (qx * qx + qz * qz);
    R[1][2] = 2 * (qy * qz - qx * qw);
    R[2][0] = 2 * (qx * qz - qy * qw);
    R[2][1] = 2 * (qy * qz + qx * qw);
    R[2][2] = 1 - 2 * (qx * qx + qy * qy);
}

// Triangulated parameter grid of one superquadric: cumulative triangle areas
// for area-proportional sampling.
struct Tess {
    int ne = 96, no = 192;
    std::vector<double> cum;  // per triangle
    std::vector<double> pts;  // (ne+1) x (no+1) x 3 local points
    double total = 0.0;
};

void tessellate(const SQ& s, Tess& t) {
    t.pts.resize(size_t(t.ne + 1) * (t.no + 1) * 3);
    for (int i = 0; i <= t.ne; i++)
        for (int j = 0; j <= t.no; j++) {
            double eta = -M_PI / 2 + M_PI * i / t.ne, om = -M_PI + 2 * M_PI * j / t.no;
            s.param(eta, om, &t.pts[(size_t(i) * (t.no + 1) + j) * 3]);
        }
    t.cum.clear();
    double acc = 0.0;
    auto P = [&](int i, int j) { return &t.pts[(size_t(i) * (t.no + 1) + j) * 3]; };
    auto area = [](const double* a, const double* b, const double* c) {
        double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
        return 0.5 * std::sqrt(x * x + y * y + z * z);
    };
    for (int i = 0; i < t.ne; i++)
        for (int j = 0; j < t.no; j++) {
            acc += area(P(i, j), P(i + 1, j), P(i + 1, j + 1));
            t.cum.push_back(acc);
            acc += area(P(i, j), P(i + 1, j + 1), P(i, j + 1));
            t.cum.push_back(acc);
        }
    t.total = acc;
}

void sample_tess(Rng& r, const SQ& s, const Tess& t, double q[3]) {
    const double u = r.uniform() * t.total;
    size_t lo = 0, hi = t.cum.size() - 1;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (t.cum[mid] < u)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int cell = int(lo / 2), half = int(lo % 2);
    const int i = cell / t.no, j = cell % t.no;
    auto P = [&](int a, int b) { return &t.pts[(size_t(a) * (t.no + 1) + b) * 3]; };
    const double* A = P(i, j);
    const double* B = half ? P(i + 1, j + 1) : P(i + 1, j);
    const double* C = half ? P(i, j + 1) : P(i + 1, j + 1);
    double a = r.uniform(), b = r.uniform();
    if (a + b > 1) {
        a = 1 - a;
        b = 1 - b;
    }
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = A[k] + a * (B[k] - A[k]) + b * (C[k] - A[k]);
    if (v[0] == 0 && v[1] == 0 && v[2] == 0) v[2] = 1e-12;
    s.project(v, q);
}

}  // namespace

extern "C" {

prg_status prg_synth_sphere(int64_t n, uint64_t seed, float* xyz, float* nxyz) {
    if (n < 1 || !xyz || !nxyz) return PRG_E_ARG;
    Rng r(seed);
    for (int64_t i = 0; i < n; i++) {
        double g[3] = {r.normal(), r.normal(), r.normal()};
        unit(g);
        for (int a = 0; a < 3; a++) {
            xyz[3 * i + a] = float(0.5 + 0.4 * g[a]);
            nxyz[3 * i + a] = float(g[a]);
        }
    }
    return PRG_OK;
}

prg_status prg_synth_torus(int64_t n, uint64_t seed, float* xyz, float* nxyz) {
    if (n < 1 || !xyz || !nxyz) return PRG_E_ARG;
    const double R = 0.30, rr = 0.12, sigma = 0.002;
    Rng r(seed);
    for (int64_t i = 0; i < n;) {
        const double th = 2 * M_PI * r.uniform(), ph = 2 * M_PI * r.uniform();
        if (r.uniform() > (R + rr * std::cos(ph)) / (R + rr)) continue;
        double nn[3] = {std::cos(ph) * std::cos(th), std::cos(ph) * std::sin(th), std::sin(ph)};
        double p[3] = {(R + rr * std::cos(ph)) * std::cos(th), (R + rr * std::cos(ph)) * std::sin(th), rr * std::sin(ph)};
        for (int a = 0; a < 3; a++) p[a] += 0.5 + sigma * r.normal();
        tilt(r, nn, 0.05);
        for (int a = 0; a < 3; a++) {
            xyz[3 * i + a] = float(p[a]);
            nxyz[3 * i + a] = float(nn[a]);
        }
        i++;
    }
    return PRG_OK;
}

prg_status prg_synth_superquadrics(int64_t n, int32_t shapes, float sigma, uint64_t seed, float* xyz, float* nxyz) {
    if (n < 1 || shapes < 1 || shapes > 16 || !xyz || !nxyz) return PRG_E_ARG;
    Rng r(seed);
    std::vector<SQ> sq(shapes);
    std::vector<Tess> ts(shapes);
    for (auto& s : sq) {
        s.e1 = r.uniform(0.3, 2.5);
        s.e2 = r.uniform(0.3, 2.5);
        for (int a = 0; a < 3; a++) s.ax[a] = r.uniform(0.1, 0.3);
        for (int a = 0; a < 3; a++) s.c[a] = r.uniform(0.3, 0.7);
        random_rotation(r, s.R);
    }
    double total = 0.0;
    std::vector<double> cum(shapes);
    for (int k = 0; k < shapes; k++) {
        tessellate(sq[k], ts[k]);
        total += ts[k].total;
        cum[k] = total;
    }
    int64_t i = 0;
    int64_t guard = 0;
    while (i < n) {
        if (++guard > 200 * n + 1000000) return PRG_E_ARG;  // shapes swallowed each other
        const double u = r.uniform() * total;
        int k = 0;
        while (k < shapes - 1 && cum[k] < u) k++;
        double q[3];
        sample_tess(r, sq[k], ts[k], q);
        // world position: p = c + R^T q
        double p[3], g[3], gw[3];
        for (int a = 0; a < 3; a++) p[a] = sq[k].c[a] + sq[k].R[0][a] * q[0] + sq[k].R[1][a] * q[1] + sq[k].R[2][a] * q[2];
        bool hidden = false;
        for (int m = 0; m < shapes && !hidden; m++)
            if (m != k && sq[m].inside_world(p)) hidden = true;
        if (hidden) continue;
        sq[k].grad(q, g);
        for (int a = 0; a < 3; a++) gw[a] = sq[k].R[0][a] * g[0] + sq[k].R[1][a] * g[1] + sq[k].R[2][a] * g[2];
        unit(gw);
        for (int a = 0; a < 3; a++) {
            xyz[3 * i + a] = float(p[a] + double(sigma) * r.normal());
            nxyz[3 * i + a] = float(gw[a]);
        }
        i++;
    }
    return PRG_OK;
}

}  // extern "C"
