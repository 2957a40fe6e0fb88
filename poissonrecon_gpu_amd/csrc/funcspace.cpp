// Host construction of the function-space tables the kernels read.
//
// The reference builds res x res double tables of B-spline inner products on
// one CPU thread (FunctionData::setDotTables, FunctionData.inl:158-215;
// O(4^(D+1)) products, 4.2 M pairs at depth 10).  Every entry is a function of
// the two depths and of the relative position (c2-c1)/w1 only: the centres and
// widths are dyadic, so the shift and scale handed to the float polynomial
// code are exact.  The kernels therefore need
//   - one 27-point Laplacian stencil per depth (GetLaplacianEntry, main.cu:1143-1158),
//   - one 1-D table of dDot weights per node depth for the divergence
//     (dot_F_DF[node + res*leaf], main.cu:1030-1054), indexed by the leaf's
//     offset u = o_leaf - (o_node - 1) * 2^(D-d) in [0, 3 * 2^(D-d)),
//   - the per-index basis polynomials baseFunctions[i] (float coefficients,
//     FunctionData.inl:141-155) and the depth-D polynomial used by the vector
//     field splat (main.cu:3362),
// computed with the same float operation sequence as the reference's
// Polynomial / PPolynomial code (Polynomial.inl, PPolynomial.inl), so the
// values are the ones the reference stores.  The piecewise polynomials here
// are fixed-size (at most 16 pieces of degree <= 4).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"

namespace prg {
namespace {

struct Pl {  // polynomial with float coefficients, c[0..deg]
    int deg = 0;
    float c[5] = {0, 0, 0, 0, 0};
};

float pl_eval(const Pl& p, float t) {  // Polynomial.inl:81-89
    float temp = 1.0f, v = 0.0f;
    for (int i = 0; i <= p.deg; i++) {
        v += temp * p.c[i];
        temp *= t;
    }
    return v;
}
float pl_integral(const Pl& p, float lo, float hi) {  // Polynomial.inl:90-101
    float v = 0.0f, t1 = lo, t2 = hi;
    for (int i = 0; i <= p.deg; i++) {
        v += p.c[i] * (t2 - t1) / float(i + 1);
        t1 *= lo;
        t2 *= hi;
    }
    return v;
}
Pl pl_deriv(const Pl& p) {  // Polynomial.inl:62-66
    Pl q;
    q.deg = p.deg - 1;
    for (int i = 0; i < p.deg; i++) q.c[i] = p.c[i + 1] * float(i + 1);
    return q;
}
Pl pl_antideriv(const Pl& p) {  // Polynomial.inl:68-73
    Pl q;
    q.deg = p.deg + 1;
    for (int i = 0; i <= p.deg; i++) q.c[i + 1] = p.c[i] / float(i + 1);
    return q;
}
Pl pl_scale(const Pl& p, float s) {  // Polynomial.inl:227-235
    Pl q = p;
    float s2 = 1.0f;
    for (int i = 0; i <= p.deg; i++) {
        q.c[i] *= s2;
        s2 /= s;
    }
    return q;
}
Pl pl_shift(const Pl& p, float t) {  // Polynomial.inl:237-248
    Pl q;
    q.deg = p.deg;
    for (int i = 0; i <= p.deg; i++) {
        float temp = 1.0f;
        for (int j = i; j >= 0; j--) {
            q.c[j] += p.c[i] * temp;
            temp *= -t * float(j);
            temp /= float(i - j + 1);
        }
    }
    return q;
}
Pl pl_mul(const Pl& a, const Pl& b) {  // Polynomial.inl:186-191
    Pl q;
    q.deg = a.deg + b.deg;
    for (int i = 0; i <= a.deg; i++)
        for (int j = 0; j <= b.deg; j++) q.c[i + j] += a.c[i] * b.c[j];
    return q;
}

struct Pc {  // piecewise polynomial: sum of pieces active for x > start
    int deg = 0;
    int n = 0;
    float start[16];
    Pl p[16];
};

// PPolynomial::set (PPolynomial.inl:84-93): stable sort by start (the
// reference's qsort is glibc's merge sort), merge equal starts in order.
Pc pc_from(int deg, int n, const float* st, const Pl* pl) {
    int order[32];
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order, order + n, [&](int a, int b) { return (st[a] - st[b]) < 0.0f; });
    Pc q;
    q.deg = deg;
    for (int k = 0; k < n; k++) {
        int i = order[k];
        if (q.n == 0 || st[i] != q.start[q.n - 1]) {
            q.start[q.n] = st[i];
            q.p[q.n] = pl[i];
            q.p[q.n].deg = deg;
            q.n++;
        } else {
            Pl& t = q.p[q.n - 1];
            for (int c = 0; c <= deg; c++) t.c[c] += pl[i].c[c];
        }
    }
    return q;
}
float pc_eval(const Pc& f, float t) {  // PPolynomial.inl:123-127
    float v = 0.0f;
    for (int i = 0; i < f.n && t > f.start[i]; i++) v += pl_eval(f.p[i], t);
    return v;
}
float pc_integral(const Pc& f, float lo, float hi) {  // PPolynomial.inl:128-143
    int m = 1;
    float a = lo, b = hi, v = 0.0f;
    if (lo > hi) {
        m = -1;
        a = hi;
        b = lo;
    }
    for (int i = 0; i < f.n && f.start[i] < b; i++) v += pl_integral(f.p[i], (a < f.start[i]) ? f.start[i] : a, b);
    return v * float(m);
}
Pc pc_scale(const Pc& f, float s) {
    Pc q = f;
    for (int i = 0; i < f.n; i++) {
        q.start[i] = f.start[i] * s;
        q.p[i] = pl_scale(f.p[i], s);
    }
    return q;
}
Pc pc_shift(const Pc& f, float t) {
    Pc q = f;
    for (int i = 0; i < f.n; i++) {
        q.start[i] = f.start[i] + t;
        q.p[i] = pl_shift(f.p[i], t);
    }
    return q;
}
Pc pc_deriv(const Pc& f) {
    Pc q = f;
    q.deg = f.deg - 1;
    for (int i = 0; i < f.n; i++) q.p[i] = pl_deriv(f.p[i]);
    return q;
}
Pc pc_mul(const Pc& a, const Pc& b) {  // PPolynomial.inl:237-250
    float st[32];
    Pl pl[32];
    int n = 0;
    for (int i = 0; i < a.n; i++)
        for (int j = 0; j < b.n; j++) {
            st[n] = (a.start[i] > b.start[j]) ? a.start[i] : b.start[j];
            pl[n] = pl_mul(a.p[i], b.p[j]);
            n++;
        }
    return pc_from(a.deg + b.deg, n, st, pl);
}
Pc pc_scalar(const Pc& f, float s, bool divide) {
    Pc q = f;
    for (int i = 0; i < f.n; i++)
        for (int c = 0; c <= f.deg; c++) q.p[i].c[c] = divide ? f.p[i].c[c] / s : f.p[i].c[c] * s;
    return q;
}
Pc pc_moving_average(const Pc& f, float radius) {  // PPolynomial.inl:372-384
    float st[32];
    Pl pl[32];
    for (int i = 0; i < f.n; i++) {
        Pl ai = pl_antideriv(f.p[i]);
        float at = pl_eval(ai, f.start[i]);
        ai.c[0] -= at;
        st[2 * i] = f.start[i] - radius;
        st[2 * i + 1] = f.start[i] + radius;
        pl[2 * i] = pl_shift(ai, -radius);
        Pl b = pl_shift(ai, radius);
        for (int c = 0; c <= b.deg; c++) b.c[c] *= -1.0f;
        pl[2 * i + 1] = b;
    }
    Pc A = pc_from(f.deg + 1, 2 * f.n, st, pl);
    return pc_scalar(pc_scalar(A, 1.0f, false), 2.0f * radius, true);
}
Pc pc_gaussian(int degree) {  // GaussianApproximation (PPolynomial.inl:349-371)
    if (degree == 0) {
        Pc q;
        q.deg = 0;
        q.n = 2;
        q.start[0] = -0.5f;
        q.start[1] = 0.5f;
        q.p[0].c[0] = 1.0f;
        q.p[1].c[0] = -1.0f;
        return q;
    }
    return pc_moving_average(pc_gaussian(degree - 1), 0.5f);
}

struct Space {
    int D;
    Pc F, dF;
    double t1, t2, r;
    explicit Space(int d) : D(d) {
        Pc G = pc_gaussian(2);
        F = pc_scalar(G, pc_eval(G, 0.0f), true);  // F / F(0), FunctionData.inl:132-134
        dF = pc_deriv(F);
        t1 = double(F.start[0]);
        t2 = double(F.start[F.n - 1]);
        r = std::fabs(double(F.start[0]));
    }
    static void cw(int depth, int64_t off, double& c, double& w) {  // BinaryNode.cuh:48-52
        w = 1.0 / double(int64_t(1) << depth);
        c = (0.5 + double(off)) * w;
    }
    // One table entry T[a + res*b] (which: 0 dot, 1 dDot, 2 d2Dot), FunctionData.inl:158-215.
    double entry(int which, int da, int64_t oa, int db, int64_t ob) const {
        int64_t ia = (int64_t(1) << da) - 1 + oa, ib = (int64_t(1) << db) - 1 + ob;
        bool a_first = (da > db) || (da == db && ia >= ib);
        double c1, w1, c2, w2;
        cw(a_first ? da : db, a_first ? oa : ob, c1, w1);
        cw(a_first ? db : da, a_first ? ob : oa, c2, w2);
        double s1 = t1 * w1 + c1, e1 = t2 * w1 + c1;
        double s = t1 * w2 + c2, e = t2 * w2 + c2;
        if (s < s1) s = s1;
        if (e > e1) e = e1;
        if (s >= e) return 0.0;
        const float sc = float(w2 / w1), sh = float((c2 - c1) / w1);
        const float lo = float(-2 * r), hi = float(2 * r);
        double dot = double(pc_integral(pc_mul(F, pc_shift(pc_scale(F, sc), sh)), lo, hi)) * w1;
        if (std::fabs(dot) < 1e-15) return 0.0;
        if (which == 0) return dot;
        if (which == 1) {
            double v = double(pc_integral(pc_mul(dF, pc_shift(pc_scale(F, sc), sh)), lo, hi));
            return (a_first && ia != ib) ? v : -v;  // the diagonal keeps -v (FunctionData.inl:201-203)
        }
        return double(pc_integral(pc_mul(dF, pc_shift(pc_scale(dF, sc), sh)), lo, hi)) / w2;
    }
};

}  // namespace

void build_fs_tables(int D, FsTables* out) {
    Space sp(D);
    out->D = D;
    out->res = (int64_t(1) << (D + 1)) - 1;
    // baseFunctions[i] = F.scale(w).shift(c) (FunctionData.inl:141-155)
    out->base_fn.assign(size_t(out->res) * 20, 0.0f);
    for (int64_t i = 0; i < out->res; i++) {
        int64_t v = i + 1;
        int depth = -1;
        while (v) {
            v >>= 1;
            depth++;
        }
        int64_t off = (i + 1) - (int64_t(1) << depth);
        double c, w;
        Space::cw(depth, off, c, w);
        Pc b = pc_shift(pc_scale(sp.F, float(w)), float(c));
        float* o = &out->base_fn[size_t(i) * 20];
        for (int k = 0; k < 4 && k < b.n; k++) {
            o[5 * k] = b.start[k];
            for (int j = 0; j <= 2; j++) o[5 * k + 1 + j] = b.p[k].c[j];
            o[5 * k + 4] = 0.0f;  // degree promoted to 3 (PPolynomial.inl:113-122)
        }
    }
    Pc bm = pc_scale(sp.F, float(1.0 / double(int64_t(1) << D)));
    std::memset(out->base_maxdepth, 0, sizeof(out->base_maxdepth));
    for (int k = 0; k < 4 && k < bm.n; k++) {
        out->base_maxdepth[4 * k] = bm.start[k];
        for (int j = 0; j <= 2; j++) out->base_maxdepth[4 * k + 1 + j] = bm.p[k].c[j];
    }
    // Laplacian stencils (GenerateSingleNodeLaplacian, main.cu:1160-1213)
    out->stencil.assign(size_t(D + 1) * 27, 0.0f);
    for (int d = 0; d <= D; d++) {
        double dt[3], d2[3];
        for (int o = -1; o <= 1; o++) {
            dt[o + 1] = sp.entry(0, d, o, d, 0);
            d2[o + 1] = sp.entry(2, d, o, d, 0);
        }
        for (int j = 0; j < 27; j++) {
            int ox = j / 9, oy = (j / 3) % 3, oz = j % 3;
            double v = d2[ox] * dt[oy] * dt[oz] + d2[oy] * dt[ox] * dt[oz] + d2[oz] * dt[ox] * dt[oy];
            out->stencil[size_t(d) * 27 + j] = (std::fabs(v) > double(float(1e-6))) ? float(v) : 0.0f;
        }
    }
    // divergence weights (computeEncodedFinerNodesDivergence, main.cu:1030-1054)
    out->div_off.assign(D + 2, 0);
    for (int d = 0; d <= D; d++) out->div_off[d + 1] = out->div_off[d] + 3 * (int64_t(1) << (D - d));
    out->div_table.assign(size_t(out->div_off[D + 1]), 0.0f);
    for (int d = 0; d <= D; d++) {
        int64_t s = int64_t(1) << (D - d);
        for (int64_t u = 0; u < 3 * s; u++)
            out->div_table[size_t(out->div_off[d] + u)] = float(sp.entry(1, d, 0, D, u - s));
    }
}

}  // namespace prg
