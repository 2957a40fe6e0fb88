// ORACLE (test infrastructure only).  CPU restatement of the reference's
// function-space arithmetic: float piecewise polynomials, the degree-2
// B-spline, the per-index scaled/shifted bases and the dot / d-dot / d2-dot
// table entries.  Only tests/, the smoke check and bench.py's CPU-baseline leg
// use this code; the product library never links it.
//
// Every routine follows the reference's float operation order literally, so a
// table entry here is the value the reference's host code stores:
//   Polynomial<Deg>            Polynomial.inl:74-249
//   StartingPolynomial / PPolynomial   PPolynomial.inl:28-389
//   GaussianApproximation      PPolynomial.inl:385-411
//   FunctionData::set / setDotTables / *DotProduct   FunctionData.inl:111-300
//   ConfirmedPPolynomial::shift / value   ConfirmedPPolynomial.cuh:37-91
//
// Pinning (the reference ships no tables and its host code needs CUDA
// declarations this image lacks): tests compare these entries with exact
// rational B-spline integrals, translation invariance, (anti)symmetry and a
// brute-force quadrature of the float polynomials.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace orc {

// ---- Polynomial<Deg> with float coefficients (Polynomial.inl) ----------------
struct Poly {
    int deg = 0;
    float c[8];
    explicit Poly(int d = 0) : deg(d) { std::memset(c, 0, sizeof(c)); }
    // Polynomial::operator() (Polynomial.inl:81-89): v += temp*c[i]; temp *= t
    float eval(float t) const {
        float temp = 1.0f, v = 0.0f;
        for (int i = 0; i <= deg; i++) {
            float term = temp * c[i];
            v = v + term;
            temp = temp * t;
        }
        return v;
    }
    // Polynomial::integral(tMin,tMax) (Polynomial.inl:90-101)
    float integral(float tMin, float tMax) const {
        float v = 0.0f, t1 = tMin, t2 = tMax;
        for (int i = 0; i <= deg; i++) {
            float a = c[i] * (t2 - t1);
            a = a / float(i + 1);
            v = v + a;
            t1 = t1 * tMin;
            t2 = t2 * tMax;
        }
        return v;
    }
    // Polynomial::derivative (Polynomial.inl:62-66)
    Poly derivative() const {
        Poly p(deg - 1);
        for (int i = 0; i < deg; i++) p.c[i] = c[i + 1] * float(i + 1);
        return p;
    }
    // Polynomial::integral() (Polynomial.inl:68-73): c'[i+1] = c[i]/(i+1)
    Poly antiderivative() const {
        Poly p(deg + 1);
        p.c[0] = 0.0f;
        for (int i = 0; i <= deg; i++) p.c[i + 1] = c[i] / float(i + 1);
        return p;
    }
    // Polynomial::scale (Polynomial.inl:227-235): c[i] *= s2; s2 /= s
    Poly scale(float s) const {
        Poly q = *this;
        float s2 = 1.0f;
        for (int i = 0; i <= deg; i++) {
            q.c[i] = q.c[i] * s2;
            s2 = s2 / s;
        }
        return q;
    }
    // Polynomial::shift (Polynomial.inl:237-248): f(x) -> f(x-t)
    Poly shift(float t) const {
        Poly q(deg);
        for (int i = 0; i <= deg; i++) {
            float temp = 1.0f;
            for (int j = i; j >= 0; j--) {
                float term = c[i] * temp;
                q.c[j] = q.c[j] + term;
                float mt = (-t) * float(j);
                temp = temp * mt;
                temp = temp / float(i - j + 1);
            }
        }
        return q;
    }
    // Polynomial * Polynomial (Polynomial.inl:186-191)
    Poly mul(const Poly& o) const {
        Poly q(deg + o.deg);
        for (int i = 0; i <= deg; i++)
            for (int j = 0; j <= o.deg; j++) {
                float term = c[i] * o.c[j];
                q.c[i + j] = q.c[i + j] + term;
            }
        return q;
    }
    Poly mul_scalar(float s) const {  // operator*(float) (Polynomial.inl:214-219)
        Poly q(deg);
        for (int i = 0; i <= deg; i++) q.c[i] = c[i] * s;
        return q;
    }
    void div_scalar(float s) {  // operator/=(float) (Polynomial.inl:207-211)
        for (int i = 0; i <= deg; i++) c[i] = c[i] / s;
    }
    void add(const Poly& o) {  // operator+= (Polynomial.inl:122-126)
        for (int i = 0; i <= deg; i++) c[i] = c[i] + o.c[i];
    }
};

struct SPoly {
    Poly p;
    float start = 0.0f;
};

// ---- PPolynomial<Deg>: sum of polynomials each active for x > start ---------
struct PPoly {
    int deg = 0;
    std::vector<SPoly> s;

    // PPolynomial::set(sps,count) (PPolynomial.inl:84-93): sort by start
    // (glibc's qsort is a merge sort here, i.e. stable), then merge equal starts
    // by adding polynomials in sorted order.
    static PPoly from_unsorted(std::vector<SPoly> sps, int deg) {
        std::stable_sort(sps.begin(), sps.end(),
                         [](const SPoly& a, const SPoly& b) { return (a.start - b.start) < 0.0f; });
        PPoly q;
        q.deg = deg;
        for (auto& sp : sps) {
            if (q.s.empty() || sp.start != q.s.back().start)
                q.s.push_back(sp);
            else
                q.s.back().p.add(sp.p);
        }
        return q;
    }
    // PPolynomial::operator() (PPolynomial.inl:123-127)
    float eval(float t) const {
        float v = 0.0f;
        for (size_t i = 0; i < s.size() && t > s[i].start; i++) v = v + s[i].p.eval(t);
        return v;
    }
    // PPolynomial::integral(tMin,tMax) (PPolynomial.inl:128-143)
    float integral(float tMin, float tMax) const {
        int m = 1;
        float start = tMin, end = tMax, v = 0.0f;
        if (tMin > tMax) {
            m = -1;
            start = tMax;
            end = tMin;
        }
        for (size_t i = 0; i < s.size() && s[i].start < end; i++) {
            float st = (start < s[i].start) ? s[i].start : start;
            v = v + s[i].p.integral(st, end);
        }
        return v * float(m);
    }
    // PPolynomial::scale (PPolynomial.inl:260-265) -> StartingPolynomial::scale
    PPoly scale(float sc) const {
        PPoly q;
        q.deg = deg;
        for (auto& sp : s) {
            SPoly o;
            o.start = sp.start * sc;
            o.p = sp.p.scale(sc);
            q.s.push_back(o);
        }
        return q;
    }
    // PPolynomial::shift (PPolynomial.inl:266-271) -> StartingPolynomial::shift
    PPoly shift(float t) const {
        PPoly q;
        q.deg = deg;
        for (auto& sp : s) {
            SPoly o;
            o.start = sp.start + t;
            o.p = sp.p.shift(t);
            q.s.push_back(o);
        }
        return q;
    }
    PPoly derivative() const {  // PPolynomial.inl:272-280
        PPoly q;
        q.deg = deg - 1;
        for (auto& sp : s) {
            SPoly o;
            o.start = sp.start;
            o.p = sp.p.derivative();
            q.s.push_back(o);
        }
        return q;
    }
    // PPolynomial * PPolynomial (PPolynomial.inl:237-250): all products, start = max
    PPoly mul(const PPoly& o) const {
        std::vector<SPoly> sp;
        for (size_t i = 0; i < s.size(); i++)
            for (size_t j = 0; j < o.s.size(); j++) {
                SPoly x;
                x.start = (s[i].start > o.s[j].start) ? s[i].start : o.s[j].start;
                x.p = s[i].p.mul(o.s[j].p);
                sp.push_back(x);
            }
        return from_unsorted(sp, deg + o.deg);
    }
    PPoly div_scalar(float v) const {  // operator/ (PPolynomial.inl:327-331)
        PPoly q = *this;
        for (auto& sp : q.s) sp.p.div_scalar(v);
        return q;
    }
    PPoly mul_scalar(float v) const {  // operator* (PPolynomial.inl:322-326)
        PPoly q = *this;
        for (auto& sp : q.s)
            for (int i = 0; i <= sp.p.deg; i++) sp.p.c[i] = sp.p.c[i] * v;
        return q;
    }
    // MovingAverage (PPolynomial.inl:372-384)
    PPoly moving_average(float radius) const {
        std::vector<SPoly> sps;
        for (auto& sp : s) {
            Poly ai = sp.p.antiderivative();
            float at = ai.eval(sp.start);
            Poly p = ai;
            p.c[0] = p.c[0] - at;  // Polynomial - float (Polynomial.inl:220-224)
            SPoly a, b;
            a.start = sp.start - radius;
            b.start = sp.start + radius;
            a.p = p.shift(-radius);
            b.p = p.shift(radius).mul_scalar(-1.0f);
            sps.push_back(a);
            sps.push_back(b);
        }
        PPoly A = from_unsorted(sps, deg + 1);
        // A*1.0/(2*radius): (A*1.0f) then / (2*radius)
        return A.mul_scalar(1.0f).div_scalar(2.0f * radius);
    }
    // ConstantFunction(radius) (PPolynomial.inl:349-361)
    static PPoly constant(float radius) {
        PPoly q;
        q.deg = 0;
        SPoly a, b;
        a.start = -radius;
        b.start = radius;
        a.p.c[0] = 1.0f;
        b.p.c[0] = -1.0f;
        q.s.push_back(a);
        q.s.push_back(b);
        return q;
    }
    // GaussianApproximation<deg>(0.5) (PPolynomial.inl:362-371)
    static PPoly gaussian(int degree) {
        if (degree == 0) return constant(0.5f);
        return gaussian(degree - 1).moving_average(0.5f);
    }
    // Degree promotion (PPolynomial<Deg+1> = PPolynomial<Deg>, PPolynomial.inl:113-122)
    PPoly promoted(int newdeg) const {
        PPoly q = *this;
        q.deg = newdeg;
        for (auto& sp : q.s) sp.p.deg = newdeg;
        return q;
    }
};

// BinaryNode<Real>::CenterAndWidth(idx) (BinaryNode.cuh:48-63)
inline void center_width_d(int depth, long long offset, double& c, double& w) {
    w = 1.0 / double(1LL << depth);
    c = (0.5 + double(offset)) * w;
}
inline void depth_offset(long long idx, int& depth, long long& offset) {
    long long i = idx + 1;
    depth = -1;
    while (i) {
        i >>= 1;
        depth++;
    }
    offset = (idx + 1) - (1LL << depth);
}

// FunctionData<2,double> with normalize = 0, useDotRatios = 0
// (main.cu:3308-3310, FunctionData.inl:111-215).
struct FunctionSpace {
    int D = 0;
    long long res = 0;
    PPoly F;       // GaussianApproximation<2>() / F(0)   (FunctionData.inl:132-134)
    PPoly dF;      // derivative
    double t1 = 0, t2 = 0, r = 0;

    explicit FunctionSpace(int maxDepth) : D(maxDepth) {
        res = (1LL << (D + 1)) - 1;
        PPoly G = PPoly::gaussian(2);
        F = G.div_scalar(G.eval(0.0f));
        dF = F.derivative();
        t1 = double(F.s.front().start);
        t2 = double(F.s.back().start);
        r = std::fabs(double(F.s.front().start));
    }
    // baseFunctions[i] = baseFunction.scale(w1).shift(c1)  (FunctionData.inl:141-155)
    PPoly base_function(long long idx) const {
        int d;
        long long o;
        depth_offset(idx, d, o);
        double c, w;
        center_width_d(d, o, c, w);
        return F.scale(float(w)).shift(float(c)).promoted(3);
    }
    // ReconstructionFunction.scale(1/2^D) as ConfirmedPPolynomial<2,4> (main.cu:3362)
    PPoly base_maxdepth() const { return F.scale(float(1.0 / double(1LL << D))); }

    // dotProduct / dDotProduct / d2DotProduct, normalize = 0 (FunctionData.inl:265-300)
    double dot_product(double c1, double w1, double c2, double w2) const {
        PPoly g = F.scale(float(w2 / w1)).shift(float((c2 - c1) / w1));
        return double(F.mul(g).integral(float(-2 * r), float(2 * r))) * w1;
    }
    double ddot_product(double c1, double w1, double c2, double w2) const {
        PPoly g = F.scale(float(w2 / w1)).shift(float((c2 - c1) / w1));
        return double(dF.mul(g).integral(float(-2 * r), float(2 * r)));
    }
    double d2dot_product(double c1, double w1, double c2, double w2) const {
        PPoly g = dF.scale(float(w2 / w1)).shift(float((c2 - c1) / w1));
        return double(dF.mul(g).integral(float(-2 * r), float(2 * r))) / w2;
    }

    // Table entry T[a + res*b] of dotTable (which=0), dDotTable (1) or
    // d2DotTable (2), following setDotTables (FunctionData.inl:158-215): the
    // pair is evaluated with the larger index first; entries whose supports do
    // not overlap or whose |dot| < 1e-15 stay 0; dDot[j + res*i] = -dDot[i + res*j],
    // which on the diagonal overwrites the entry with its negative (:201-203).
    // Indices are given as (depth, offset) so offsets outside [0, 2^depth) work too.
    double entry(int which, int da, long long oa, int db, long long ob) const {
        long long ia = (1LL << da) - 1 + oa, ib = (1LL << db) - 1 + ob;
        bool a_first = (da > db) || (da == db && ia >= ib);
        int d1 = a_first ? da : db, d2 = a_first ? db : da;
        long long o1 = a_first ? oa : ob, o2 = a_first ? ob : oa;
        double c1, w1, c2, w2;
        center_width_d(d1, o1, c1, w1);
        center_width_d(d2, o2, c2, w2);
        double start1 = t1 * w1 + c1, end1 = t2 * w1 + c1;
        double start = t1 * w2 + c2, end = t2 * w2 + c2;
        if (start < start1) start = start1;
        if (end > end1) end = end1;
        if (start >= end) return 0.0;
        double dot = dot_product(c1, w1, c2, w2);
        if (std::fabs(dot) < 1e-15) return 0.0;
        if (which == 0) return dot;
        if (which == 1) {
            // idx1 = i + res*j gets +v and idx2 = j + res*i gets -v; on the
            // diagonal idx1 == idx2, so the second store leaves -v.
            double v = ddot_product(c1, w1, c2, w2);
            return (a_first && ia != ib) ? v : -v;
        }
        return d2dot_product(c1, w1, c2, w2);
    }
    double entry_idx(int which, long long a, long long b) const {
        int da, db;
        long long oa, ob;
        depth_offset(a, da, oa);
        depth_offset(b, db, ob);
        return entry(which, da, oa, db, ob);
    }
};

// ConfirmedPPolynomial<Deg,4>::shift (ConfirmedPPolynomial.cuh:37-51) on the
// first `count` pieces: coefficients accumulate into zeroed storage.
struct CPoly {
    int deg;
    float start[4];
    float c[4][4];
};
inline CPoly cpoly_from(const PPoly& p) {
    CPoly q;
    q.deg = p.deg;
    std::memset(q.start, 0, sizeof(q.start));
    std::memset(q.c, 0, sizeof(q.c));
    for (int i = 0; i < 4 && i < int(p.s.size()); i++) {
        q.start[i] = p.s[i].start;
        for (int j = 0; j <= p.deg; j++) q.c[i][j] = p.s[i].p.c[j];
    }
    return q;
}
inline CPoly cpoly_shift(const CPoly& p, float t) {
    CPoly q;
    q.deg = p.deg;
    std::memset(q.c, 0, sizeof(q.c));
    for (int i = 0; i < 4; i++) {
        q.start[i] = p.start[i] + t;
        for (int k = 0; k <= p.deg; k++) {
            float temp = 1.0f;
            for (int j = k; j >= 0; j--) {
                float term = p.c[i][k] * temp;
                q.c[i][j] = q.c[i][j] + term;
                float mt = (-t) * float(j);
                temp = temp * mt;
                temp = temp / float(k - j + 1);
            }
        }
    }
    return q;
}
// value() (ConfirmedPPolynomial.cuh:78-91)
inline float cpoly_value(const CPoly& p, float x) {
    float res = 0.0f;
    for (int i = 0; i < 4 && x > p.start[i]; i++) {
        float temp = 1.0f, v = 0.0f;
        for (int j = 0; j <= p.deg; j++) {
            float term = temp * p.c[i][j];
            v = v + term;
            temp = temp * x;
        }
        res = res + v;
    }
    return res;
}

}  // namespace orc
