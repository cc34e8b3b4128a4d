// Numeric core shared by the host loader and the gfx950 kernels.
//
// Double-double arithmetic follows src/math/dfloat_t.h:83-175 (TwoSum,
// TwoProd with a true FMA, quick renormalisation); vector and matrix algebra
// follows src/math/vec_t.h.  Everything is compiled with -ffp-contract=off,
// so each operator is exactly one IEEE operation in the written order:
// vector division multiplies by a reciprocal, three-term sums fold left,
// matrices are stored as three axis columns.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define EPH_HD __host__ __device__ __forceinline__

namespace eph {

struct dd { double hi, lo; };
struct v3 { double x, y, z; };
struct m3 { v3 x, y, z; };
struct ddv3 { dd x, y, z; };
struct ddm3 { ddv3 x, y, z; };

constexpr double kC = 299792458.0;
constexpr double kC2 = kC * kC;
constexpr double kPi = 3.1415926535897932;
constexpr double kG = 6.67430E-11;

// ---- double-double --------------------------------------------------------
EPH_HD dd mkdd(double h) { return dd{h, 0.0}; }
EPH_HD double todbl(dd a) { return a.hi + a.lo; }
EPH_HD dd add(dd x, dd y) {
    double r = x.hi + y.hi, t = r - x.hi;
    double e = (x.hi - (r - t)) + (y.hi - t);
    e = e + (x.lo + y.lo);
    double h = r + e;
    return dd{h, e - (h - r)};
}
EPH_HD dd sub(dd x, dd y) {
    double r = x.hi - y.hi, t = r - x.hi;
    double e = (x.hi - (r - t)) - (y.hi + t);
    e = e + (x.lo - y.lo);
    double h = r + e;
    return dd{h, e - (h - r)};
}
EPH_HD dd mul(dd x, dd y) {
    double r = x.hi * y.hi;
    double e = fma(x.hi, y.hi, -r);
    e = e + (x.hi * y.lo + y.hi * x.lo);
    double h = r + e;
    return dd{h, e - (h - r)};
}
EPH_HD dd div(dd x, dd y) {
    double h = x.hi / y.hi;
    double sh = h * y.hi;
    double sl = fma(h, y.hi, -sh);
    double l = (x.hi - sh - sl + x.lo - h * y.lo) / y.hi;
    double hh = h + l;
    return dd{hh, l - (hh - h)};
}
EPH_HD bool finite(dd a) { return !((0 * a.hi) != 0 || (0 * a.lo) != 0); }

// ---- vectors ----------------------------------------------------------------
EPH_HD v3 mkv(double x, double y, double z) { return v3{x, y, z}; }
EPH_HD v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
EPH_HD v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
EPH_HD v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
EPH_HD v3 operator*(double s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
EPH_HD v3 operator*(v3 a, double s) { return v3{s * a.x, s * a.y, s * a.z}; }
EPH_HD v3 operator/(v3 a, double s) { double r = 1 / s; return v3{a.x * r, a.y * r, a.z * r}; }
EPH_HD v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
EPH_HD double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
EPH_HD v3 &operator+=(v3 &a, v3 b) { a = a + b; return a; }
EPH_HD v3 &operator-=(v3 &a, v3 b) { a = a - b; return a; }

// hypot_t<double>::_normalize (vec_t.h:53-81): scaled for tiny/huge norms;
// the zero vector becomes (0,0,1) for three arguments, (0,1) for two.
constexpr double kHypHigh = 0x1p960, kHypMul = 0x1p768;
constexpr double kHypDiv = 1 / kHypMul, kHypLow = 1 / kHypHigh;
EPH_HD double normalize(v3 &a) {
    double x = a.x, y = a.y, z = a.z;
    double r = x * x + y * y + z * z;
    double f = 1;
    if (r < kHypLow) {
        x *= kHypMul; y *= kHypMul; z *= kHypMul;
        r = x * x + y * y + z * z;
        if (r == 0) { a = v3{0, 0, 1}; return r; }
        f = kHypDiv;
    } else if (!(r < kHypHigh)) {
        x *= kHypDiv; y *= kHypDiv; z *= kHypDiv;
        r = x * x + y * y + z * z;
        f = kHypMul;
    }
    double s = sqrt(r), rs = 1 / s;
    a = v3{x * rs, y * rs, z * rs};
    return f == 1 ? s : f * s;
}
EPH_HD double normalize2(double &pa, double &pb) {
    double a = pa, b = pb;
    double r = a * a + b * b;
    double f = 1;
    if (r < kHypLow) {
        a *= kHypMul; b *= kHypMul;
        r = a * a + b * b;
        if (r == 0) { pa = 0; pb = 1; return r; }
        f = kHypDiv;
    } else if (!(r < kHypHigh)) {
        a *= kHypDiv; b *= kHypDiv;
        r = a * a + b * b;
        f = kHypMul;
    }
    double s = sqrt(r), rs = 1 / s;
    pa = a * rs; pb = b * rs;
    return f == 1 ? s : f * s;
}
EPH_HD v3 unit(v3 a) { normalize(a); return a; }
// hypot_t<double>::length (vec_t.h:38-50): the norm, scaled for tiny/huge
EPH_HD double length(v3 a) {
    double x = a.x, y = a.y, z = a.z;
    double r = x * x + y * y + z * z;
    if (r < kHypLow) {
        x *= kHypMul; y *= kHypMul; z *= kHypMul;
        r = x * x + y * y + z * z;
        return kHypDiv * sqrt(r);
    }
    if (!(r < kHypHigh)) {
        x *= kHypDiv; y *= kHypDiv; z *= kHypDiv;
        r = x * x + y * y + z * z;
        return kHypMul * sqrt(r);
    }
    return sqrt(r);
}
EPH_HD v3 asc_node(v3 a) {                         // vec_t.h:226-230
    v3 p{-a.y, a.x, 0};
    normalize2(p.y, p.x);
    return p;
}

// ---- matrices (three axis columns) -------------------------------------------
EPH_HD m3 mzero() { return m3{v3{0, 0, 0}, v3{0, 0, 0}, v3{0, 0, 0}}; }
EPH_HD v3 operator*(const m3 &a, v3 b) {            // mat % vec
    return (b.x * a.x + b.y * a.y) + b.z * a.z;
}
EPH_HD v3 tolocal(const m3 &s, v3 a) { return v3{dot(a, s.x), dot(a, s.y), dot(a, s.z)}; }
EPH_HD m3 operator*(const m3 &a, const m3 &b) { return m3{a * b.x, a * b.y, a * b.z}; }
EPH_HD m3 outer(v3 a, v3 b) { return m3{b.x * a, b.y * a, b.z * a}; }
EPH_HD m3 operator*(double s, const m3 &a) { return m3{s * a.x, s * a.y, s * a.z}; }
EPH_HD m3 operator+(const m3 &a, const m3 &b) { return m3{a.x + b.x, a.y + b.y, a.z + b.z}; }
EPH_HD m3 operator-(const m3 &a, const m3 &b) { return m3{a.x - b.x, a.y - b.y, a.z - b.z}; }
EPH_HD m3 operator-(const m3 &a) { return m3{-a.x, -a.y, -a.z}; }
EPH_HD m3 adddiag(m3 a, double s) { a.x.x += s; a.y.y += s; a.z.z += s; return a; }
EPH_HD m3 subdiag(m3 a, double s) { a.x.x -= s; a.y.y -= s; a.z.z -= s; return a; }
EPH_HD m3 transpose(const m3 &a) {
    return m3{v3{a.x.x, a.y.x, a.z.x}, v3{a.x.y, a.y.y, a.z.y}, v3{a.x.z, a.y.z, a.z.z}};
}
EPH_HD double det(const m3 &a) { return dot(cross(a.x, a.y), a.z); }
EPH_HD m3 inverse(const m3 &m) {                    // vec_t.h:443-454
    m3 a{v3{m.y.y * m.z.z - m.y.z * m.z.y, m.z.y * m.x.z - m.x.y * m.z.z, m.x.y * m.y.z - m.x.z * m.y.y},
         v3{m.y.z * m.z.x - m.y.x * m.z.z, m.z.z * m.x.x - m.x.z * m.z.x, m.x.z * m.y.x - m.x.x * m.y.z},
         v3{m.y.x * m.z.y - m.y.y * m.z.x, m.z.x * m.x.y - m.x.x * m.z.y, m.x.x * m.y.y - m.x.y * m.y.x}};
    double rdet = 1 / (a.x.z * m.z.x + a.y.z * m.z.y + a.z.z * m.z.z);
    return rdet * a;
}
EPH_HD m3 toworld(const m3 &s, const m3 &a) {      // vec_t.h:335-338: s a s^T
    m3 sa = s * a;
    return m3{(s.x.x * sa.x + s.y.x * sa.y) + s.z.x * sa.z,
              (s.x.y * sa.x + s.y.y * sa.y) + s.z.y * sa.z,
              (s.x.z * sa.x + s.y.z * sa.y) + s.z.z * sa.z};
}
EPH_HD m3 tolocal(const m3 &s, const m3 &a) {      // vec_t.h:330-333: s^T a s
    m3 sta{tolocal(s, a.x), tolocal(s, a.y), tolocal(s, a.z)};
    return sta * s;
}
EPH_HD m3 from_harmonics(double j2, double c21, double c22, double s21, double s22) {
    m3 m;                                           // vec_t.h:522-530
    m.x.x = 3 * c22 + j2 / 2;
    m.y.x = m.x.y = 3 * s22;
    m.z.x = m.x.z = 3 * c21 / 2;
    m.y.y = -3 * c22 + j2 / 2;
    m.z.y = m.y.z = 3 * s21 / 2;
    m.z.z = -j2;
    return m;
}
// (I+R) rotates about w by |w| t (vec_t.h:540-559).  SIN is the sine used
// for the two angle evaluations.
template <typename SinF>
EPH_HD m3 rotation_matrix(v3 w, double t, SinF sinf_) {
    v3 u = w;
    double wt = normalize(u) * t;
    double st = sinf_(wt), st222 = sinf_(wt / 2);
    st222 *= 2 * st222;
    double x = u.x, y = u.y, z = u.z;
    m3 dr{v3{x * x - 1, x * y, x * z}, v3{x * y, y * y - 1, y * z}, v3{x * z, y * z, z * z - 1}};
    dr = st222 * dr;
    dr.x.y += z * st;
    dr.y.x -= z * st;
    dr.x.z -= y * st;
    dr.z.x += y * st;
    dr.y.z += x * st;
    dr.z.y -= x * st;
    return dr;
}

// ---- double-double vectors/matrices -------------------------------------------
EPH_HD ddv3 mkddv(v3 a) { return ddv3{mkdd(a.x), mkdd(a.y), mkdd(a.z)}; }
EPH_HD v3 todbl(ddv3 a) { return v3{todbl(a.x), todbl(a.y), todbl(a.z)}; }
EPH_HD ddv3 add(ddv3 a, ddv3 b) { return ddv3{add(a.x, b.x), add(a.y, b.y), add(a.z, b.z)}; }
EPH_HD v3 diff(ddv3 a, ddv3 b) {                   // fast_mpvec(a - b)
    return v3{todbl(sub(a.x, b.x)), todbl(sub(a.y, b.y)), todbl(sub(a.z, b.z))};
}
EPH_HD ddv3 scale(dd s, ddv3 a) { return ddv3{mul(s, a.x), mul(s, a.y), mul(s, a.z)}; }
EPH_HD ddv3 sub(ddv3 a, ddv3 b) { return ddv3{sub(a.x, b.x), sub(a.y, b.y), sub(a.z, b.z)}; }
// vec_t<dfloat_t> cross product (vec_t.h), every operation double-double
EPH_HD ddv3 cross(ddv3 a, ddv3 b) {
    return ddv3{sub(mul(a.y, b.z), mul(a.z, b.y)), sub(mul(a.z, b.x), mul(a.x, b.z)), sub(mul(a.x, b.y), mul(a.y, b.x))};
}
// vec_t<dfloat_t> / dfloat_t (vec_t.h:262-265): multiply by the reciprocal
EPH_HD ddv3 divide(ddv3 a, dd s) { const dd rs = div(mkdd(1), s); return scale(rs, a); }
EPH_HD bool finite(ddv3 a) { return finite(a.z) && finite(a.y) && finite(a.x); }
EPH_HD m3 todbl(const ddm3 &a) { return m3{todbl(a.x), todbl(a.y), todbl(a.z)}; }
EPH_HD ddm3 mkddm(const m3 &a) { return ddm3{mkddv(a.x), mkddv(a.y), mkddv(a.z)}; }
EPH_HD ddm3 add(const ddm3 &a, const m3 &b) {
    return ddm3{add(a.x, mkddv(b.x)), add(a.y, mkddv(b.y)), add(a.z, mkddv(b.z))};
}
// s += 0.5 * s (1 - s^T s)  (mass.h:147-150)
EPH_HD ddm3 orthogonalized(const ddm3 &s) {
    m3 f = todbl(s);
    m3 t = adddiag(-(transpose(f) * f), 1);
    return add(s, (0.5 * f) * t);
}

EPH_HD double smax(double a, double b) { return (a < b) ? b : a; }           // std::max
EPH_HD double checked_max_into(double cur, double b) {                         // basic_math.h:34-46
    return (cur != cur || b < cur) ? cur : b;
}

}  // namespace eph
