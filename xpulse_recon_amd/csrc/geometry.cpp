// Cone-beam geometry: per-(angle, module) ProjData records (Geometry,
// dataset/geometry.hpp:35-328).  One templated body serves the float32
// construction (the reference builds its matrices in float32 with glm, so this
// mode follows its operation order) and the FP64 construction (rounded to
// float32 only when stored).  Host code; compiled with -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/xpr_hip.h"

namespace xpr_geom {

template <typename T> struct V3 {
    T x, y, z;
};
template <typename T> V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
// glm::dot: component products summed left to right
template <typename T> T dot(V3<T> a, V3<T> b)
{
    const T tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
    return (tx + ty) + tz;
}
template <typename T> V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <typename T> V3<T> normalize(V3<T> a) { return a * (T(1) / std::sqrt(dot(a, a))); }
template <typename T> T length(V3<T> a) { return std::sqrt(dot(a, a)); }

template <typename T> T radians(T deg) { return deg * T(0.01745329251994329576923690768489); }
template <typename T> T degrees(T rad) { return rad * T(57.295779513082320876798154814105); }

// rotations about the frame axes (glm rotateX / rotateY / rotateZ)
template <typename T> V3<T> rot_x(V3<T> v, T a)
{
    const T c = std::cos(a), s = std::sin(a);
    return {v.x, v.y * c - v.z * s, v.y * s + v.z * c};
}
template <typename T> V3<T> rot_y(V3<T> v, T a)
{
    const T c = std::cos(a), s = std::sin(a);
    return {v.x * c + v.z * s, v.y, -v.x * s + v.z * c};
}
template <typename T> V3<T> rot_z(V3<T> v, T a)
{
    const T c = std::cos(a), s = std::sin(a);
    return {v.x * c - v.y * s, v.x * s + v.y * c, v.z};
}
// Rodrigues rotation of v by angle a about the unit axis k
template <typename T> V3<T> rot_axis(V3<T> v, T a, V3<T> k)
{
    const T c = std::cos(a), s = std::sin(a);
    const V3<T> kxv = cross(k, v);
    return v * c + kxv * s + k * (dot(k, v) * (T(1) - c));
}

// Geometry::rotate(vec, origin, 0, deg, 0) (geometry.hpp:88-95)
template <typename T> V3<T> rotate_y_about(V3<T> vec, V3<T> origin, T deg)
{
    V3<T> v = vec - origin;
    v = rot_x(v, radians(T(0)));
    v = rot_y(v, radians(deg));
    v = rot_z(v, radians(T(0)));
    return v + origin;
}

// 4x4 as glm stores it: m[col][row]; product column c = sum_k a[k] * b[c][k]
template <typename T> struct M4 {
    T m[4][4];
};
template <typename T> M4<T> mul(const M4<T> &a, const M4<T> &b)
{
    M4<T> r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            T acc = a.m[0][row] * b.m[c][0];
            acc = acc + a.m[1][row] * b.m[c][1];
            acc = acc + a.m[2][row] * b.m[c][2];
            acc = acc + a.m[3][row] * b.m[c][3];
            r.m[c][row] = acc;
        }
    return r;
}
// glm::mat4(16 scalars) is filled column by column
template <typename T> M4<T> cols(const T (&v)[16])
{
    M4<T> r;
    for (int i = 0; i < 16; ++i) r.m[i / 4][i % 4] = v[i];
    return r;
}

// Off-axis perspective of a planar detector seen from pe, mapped to viewport
// pixels (computeGeneralizedProjection, geometry.hpp:182-234).
template <typename T>
M4<T> generalized_projection(V3<T> pe, V3<T> pa, V3<T> pb, V3<T> pc, T vpl, T vpr, T vpb,
                             T vpt, T sd)
{
    const V3<T> vr = normalize(pb - pa);
    const V3<T> vu = normalize(pc - pa);
    const V3<T> vn = normalize(cross(vu, vr));
    const V3<T> va = pa - pe, vb = pb - pe, vc = pc - pe;
    const T d = -dot(va, vn), n = T(1), f = sd * T(2);
    const T l = dot(vr, va) * (n / d), r = dot(vr, vb) * (n / d);
    const T b = dot(vu, va) * (n / d), t = dot(vu, vc) * (n / d);
    const T z = T(0), one = T(1), two = T(2), half = T(0.5);
    const T P[16] = {(two * n) / (r - l), z, (r + l) / (r - l), z,
                     z, (two * n) / (t - b), (t + b) / (t - b), z,
                     z, z, -(f + n) / (f - n), -(two * f * n) / (f - n),
                     z, z, -one, z};
    const T Mt[16] = {vr.x, vr.y, vr.z, z, vu.x, vu.y, vu.z, z, vn.x, vn.y, vn.z, z, z, z, z, one};
    const T Tr[16] = {one, z, z, -pe.x, z, one, z, -pe.y, z, z, one, -pe.z, z, z, z, one};
    const T S[16] = {(vpr - vpl) * half, z, z, (vpr + vpl) * half,
                     z, (vpt - vpb) * half, z, (vpt + vpb) * half,
                     z, z, half, half, z, z, z, one};
    return mul(mul(mul(cols(Tr), cols(Mt)), cols(P)), cols(S));
}

// frames: optional, 14 doubles per record: source pe, upper-left corner pa,
// column step er = (pb-pa)/(vpr-vpl), row step eu = (pc-pa)/(vpt-vpb), vpl, vpb
template <typename T>
void build(const xpr_geometry_params &p, float *out, xpr_grid &grid, double *frames = nullptr)
{
    const T so = T(p.so), sd = T(p.sd), px = T(p.px), sx = T(p.sx), sy = T(p.sy);
    const T rx = T(p.rx) * px;                                   // parameters.hpp:39
    const T vx = px * (so / sd);                                 // geometry.hpp:50
    const int64_t vwidth = p.dwidth;                             // :55-56
    const int64_t vheight = p.dheight + int64_t(std::fabs(T(p.heli_offset) / vx));
    // source frame, then object frame (:58-68)
    V3<T> src{T(0), T(0), T(0)};
    V3<T> at = src + V3<T>{T(0), T(0), sd};
    const V3<T> axis = at + V3<T>{-sx + rx, -sy, T(0)};
    const V3<T> oc = src + normalize(axis - src) * so;
    src = src - oc;
    at = at - oc;
    const V3<T> zero{T(0), T(0), T(0)};
    grid.vx = double(vx);
    grid.vwidth = vwidth;
    grid.vheight = vheight;
    grid.orig[0] = double(T(-0.5) * T(vwidth) * vx);
    grid.orig[1] = double(T(-0.5) * T(vheight) * vx);
    grid.orig[2] = double(T(-0.5) * T(vwidth) * vx);

    const int64_t nmod = p.module_number;
    const uint64_t width = uint64_t(p.dwidth) / uint64_t(nmod);
    const uint64_t height = uint64_t(p.dheight);
    const T heli_off = T(p.heli_offset), angle = T(p.angle);
    const bool module_rot = p.roll != 0.0 || p.pitch != 0.0 || p.yaw != 0.0;
    for (int64_t a = 0; a < p.projections; ++a) {
        const T adeg = p.angle_list ? T(p.angle_list[a])
                                    : T(a) * (angle / T(p.projections));     // :41-44
        // computeDetectorProjection (:267-291)
        const T heli = (heli_off / T(p.projections)) * T(a) - heli_off * T(0.5);
        const V3<T> pos = rotate_y_about(src, zero, adeg) + V3<T>{T(0), heli, T(0)};
        const V3<T> tgt = rotate_y_about(at, zero, adeg) + V3<T>{T(0), heli, T(0)};
        for (int64_t m = 0; m < nmod; ++m) {
            // detectorFromSource (:244-265)
            const V3<T> vdir{T(0), T(1), T(0)};
            const V3<T> u = normalize(cross(vdir, tgt - pos));
            const V3<T> det = (tgt - u * sx) - vdir * sy;
            V3<T> ul = (det - u * (T(0.5) * T(width) * px)) - vdir * (T(0.5) * T(height) * px);
            V3<T> ur = ul + u * (T(width) * px);
            V3<T> bl = ul + vdir * (T(height) * px);
            if (module_rot) {
                // module rotation about the detector centre: yaw about u, pitch
                // about v, roll about the beam axis (the formula commented out at
                // geometry.hpp:260-262)
                const V3<T> nrm = normalize(tgt - pos);
                for (V3<T> *c : {&ul, &ur, &bl}) {
                    V3<T> w = *c - det;
                    w = rot_axis(w, radians(T(p.yaw)), u);
                    w = rot_axis(w, radians(T(p.pitch)), vdir);
                    w = rot_axis(w, radians(T(p.roll)), nrm);
                    *c = w + det;
                }
            }
            const V3<T> mc = det + normalize(pos - tgt) * (sd - T(p.module_center_offset_z));
            T mang;
            if (p.has_module_angle)
                mang = T(p.module_angle_offset);
            else
                mang = degrees(T(2) * std::atan2(T(width) * T(0.5) * px, length(det - mc)));
            const T cur = mang * (T(m) - std::floor(T(nmod) / T(2)));
            ul = rotate_y_about(ul, mc, cur);
            ur = rotate_y_about(ur, mc, cur);
            bl = rotate_y_about(bl, mc, cur);
            const T vpl = T(width * uint64_t(m)), vpr = T(width * uint64_t(m + 1));
            const T vpt = T(p.dheight);
            if (frames) {
                double *f = frames + 14 * (a * nmod + m);
                const V3<T> er = (ur - ul) * (T(1) / (vpr - vpl));
                const V3<T> eu = (bl - ul) * (T(1) / vpt);
                const V3<T> pts[4] = {pos, ul, er, eu};
                for (int i = 0; i < 4; ++i) {
                    f[3 * i + 0] = double(pts[i].x);
                    f[3 * i + 1] = double(pts[i].y);
                    f[3 * i + 2] = double(pts[i].z);
                }
                f[12] = double(vpl);
                f[13] = 0.0;
            }
            if (!out) continue;
            const M4<T> M = generalized_projection(pos, ul, ur, bl, vpl, vpr, T(0), vpt, sd);
            float *o = out + 20 * (a * nmod + m);
            for (int i = 0; i < 16; ++i) o[i] = float(M.m[i / 4][i % 4]);
            o[16] = float(vpl);
            o[17] = float(vpr);
            o[18] = 0.0f;
            o[19] = float(vpt);
        }
    }
}

} // namespace xpr_geom

// ---------------------------------------------------------------------------
// Voxel hits per detector pixel of one forward launch (no counterpart in the
// reference: it lets the engine use 32-bit detector sums exactly).
//
// H bounds the number of voxels of the launch whose computed nearest pixel is
// one given pixel of one view, for every origin in [origin, origin + voxel]
// per axis (any jitter).  Every detector sum is then at most H times the
// largest Q.22 voxel value.
//
// Voxel i = (x, layer, depth) sits at P = origin + voxel * i, inside the box
// from origin + voxel * vmin to origin + voxel * vmax.  The projection
// u = X/W, v = Y/W (ProjData rows 0, 1 and 3) is smooth where W > 0, which
// must hold on the box.  The projector's float evaluation (the rounded
// position, three FMAs per row, the reciprocal, the product) stays within
//   du = 24 eps ((Tx + U Tw) / Wmin + U),  eps = 2^-24,
// of the exact u of the unrounded position (three times the constant of the
// forward's tile bound; Tx, Tw: sums of |coefficient x coordinate|, U the
// largest |u|), likewise dv, so two voxels on one pixel have exact u within
// 1 + 2 du and v within 1 + 2 dv of each other.  In one lattice plane (x,
// layer or depth fixed) their difference d satisfies, by the mean value
// theorem, du_diff = grad u(xi) . d and dv_diff = grad v(eta) . d with every
// gradient entry in an interval over the box: d u / d i_a =
// voxel_a (m_a W - m_(12+a) X) / W^2, whose affine numerator takes its range
// at the corners and W^2 between its corner extremes.  With those rows
// scaled by 1 / (1 + 2 du), 1 / (1 + 2 dv), diagonal dominance of either
// pairing of rows and in-plane axes gives |J d|_inf >= mu |d|_inf for every
// J of the intervals, hence |d|_inf <= D = 1 / mu, and each row then bounds
// its dominant axis's span: |J_ux| |d_x| <= 1 + |J_uy| D.  The voxels of one
// plane on one pixel span at most floor(span) + 1 lattice values per
// in-plane axis.  A record's bound is the least over the three plane
// families and two pairings of (planes) x (product of the two counts); a
// view's is the sum over its modules.
// ---------------------------------------------------------------------------
namespace {

struct Interval {
    double lo, hi;
};
double abs_lo(Interval a) { return a.lo > 0.0 ? a.lo : a.hi < 0.0 ? -a.hi : 0.0; }
double abs_hi(Interval a) { return std::fmax(std::fabs(a.lo), std::fabs(a.hi)); }

double record_hits(const float *m, const double (&lo)[3], const double (&hi)[3],
                   const double (&vox)[3], const int64_t (&n)[3], double line_off)
{
    const double inf = HUGE_VAL;
    double X[8], Y[8], W[8], tx = 0.0, ty = 0.0, tw = 0.0, xmax = 0.0, ymax = 0.0;
    double wlo = inf, whi = -inf;
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
        X[c] = double(m[0]) * p[0] + double(m[1]) * p[1] + double(m[2]) * p[2] + double(m[3]);
        Y[c] = double(m[4]) * p[0] + double(m[5]) * p[1] + double(m[6]) * p[2] + double(m[7]);
        W[c] = double(m[12]) * p[0] + double(m[13]) * p[1] + double(m[14]) * p[2] + double(m[15]);
        double ax = std::fabs(double(m[3])), ay = std::fabs(double(m[7])), aw = std::fabs(double(m[15]));
        for (int a = 0; a < 3; ++a) {
            ax += std::fabs(double(m[a]) * p[a]);
            ay += std::fabs(double(m[4 + a]) * p[a]);
            aw += std::fabs(double(m[12 + a]) * p[a]);
        }
        tx = std::fmax(tx, ax);
        ty = std::fmax(ty, ay);
        tw = std::fmax(tw, aw);
        xmax = std::fmax(xmax, std::fabs(X[c]));
        ymax = std::fmax(ymax, std::fabs(Y[c]));
        wlo = std::fmin(wlo, W[c]);
        whi = std::fmax(whi, W[c]);
    }
    if (!(wlo > 0.0) || !std::isfinite(whi) || !std::isfinite(tx + ty + tw)) return inf;
    constexpr double eps = 1.0 / 16777216.0;
    const double U = xmax / wlo + 1.0, V = ymax / wlo + std::fabs(line_off) + 1.0;
    const double su = 1.0 + 2.0 * (24.0 * eps * ((tx + U * tw) / wlo + U) + 1e-4);
    const double sv = 1.0 + 2.0 * (24.0 * eps * ((ty + V * tw) / wlo + V) + 1e-4);
    // g[r][a]: interval of d(u, v)[r] / d i_a over the box, scaled by 1 / s
    Interval g[2][3];
    for (int r = 0; r < 2; ++r) {
        const double *XY = r ? Y : X;
        for (int a = 0; a < 3; ++a) {
            double nlo = inf, nhi = -inf, mag = 0.0;
            for (int c = 0; c < 8; ++c) {
                const double t0 = double(m[4 * r + a]) * W[c], t1 = double(m[12 + a]) * XY[c];
                nlo = std::fmin(nlo, t0 - t1);
                nhi = std::fmax(nhi, t0 - t1);
                mag = std::fmax(mag, std::fabs(t0) + std::fabs(t1));
            }
            nlo -= 1e-12 * mag;                          // the double evaluation above
            nhi += 1e-12 * mag;
            const double w2lo = wlo * wlo, w2hi = whi * whi;
            Interval q;
            if (nlo >= 0.0) q = {nlo / w2hi, nhi / w2lo};
            else if (nhi <= 0.0) q = {nlo / w2lo, nhi / w2hi};
            else q = {nlo / w2lo, nhi / w2lo};
            const double k = vox[a] / (r ? sv : su);
            g[r][a] = {q.lo * k, q.hi * k};
        }
    }
    // voxels of one plane on one pixel when row p of the scaled gradient is
    // dominated by in-plane axis x and row q by axis y: |d|inf <= D = 1 / mu,
    // then each row alone, |J_px| |d_x| <= 1 + |J_py| |d_y| <= 1 + |J_py| D,
    // bounds that axis's span (likewise y); floor(span) + 1 lattice values
    auto plane_count = [&](const Interval &px, const Interval &py, const Interval &qx,
                           const Interval &qy) {
        const double mu = std::fmin(abs_lo(px) - abs_hi(py), abs_lo(qy) - abs_hi(qx)) * (1.0 - 1e-9);
        if (!(mu > 0.0)) return inf;
        const double D = 1.0 / mu;
        const double sx = std::fmin(D, (1.0 + abs_hi(py) * D) / abs_lo(px) * (1.0 + 1e-9));
        const double sy = std::fmin(D, (1.0 + abs_hi(qx) * D) / abs_lo(qy) * (1.0 + 1e-9));
        return (std::floor(sx) + 1.0) * (std::floor(sy) + 1.0);
    };
    double best = inf;
    for (int c = 0; c < 3; ++c) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        const double k = std::fmin(plane_count(g[0][a], g[0][b], g[1][a], g[1][b]),   // u by a, v by b
                                   plane_count(g[0][b], g[0][a], g[1][b], g[1][a]));  // u by b, v by a
        best = std::fmin(best, double(n[c]) * k);
    }
    return best;
}

} // namespace

// error reporting shared with the kernels' translation unit
int xprint_fail(int code, const char *msg);

extern "C" int xpr_hit_bound(const float *projdata, const xpr_proj_args *a, double *hits)
{
    if (!projdata || !a || !hits) return xprint_fail(XPR_E_INVALID, "null pointer");
    if (a->view_index) return xprint_fail(XPR_E_INVALID, "xpr_hit_bound: indexed views");
    double lo[3], hi[3], vox[3];
    int64_t n[3];
    for (int i = 0; i < 3; ++i) {
        if (a->vmax[i] < a->vmin[i]) return xprint_fail(XPR_E_INVALID, "xpr_hit_bound: empty box");
        n[i] = int64_t(a->vmax[i]) - a->vmin[i];
        vox[i] = double(a->voxel[i]);
        if (!(vox[i] > 0.0)) return xprint_fail(XPR_E_INVALID, "xpr_hit_bound: voxel size");
        const double o = double(a->origin[i]);
        lo[i] = o + vox[i] * a->vmin[i];
        hi[i] = o + vox[i] * a->vmax[i];
        // the origin may carry any jitter in [0, voxel] after its float rounding
        lo[i] -= 1e-6 * std::fabs(lo[i]);
        hi[i] += 1e-6 * std::fabs(hi[i]);
    }
    double worst = 0.0;
    if (n[0] > 0 && n[1] > 0 && n[2] > 0) {
        const uint32_t mods = a->modules ? a->modules : 1;
        for (uint32_t v = 0; v < a->n_views; ++v) {
            double h = 0.0;
            for (uint32_t k = 0; k < mods; ++k)
                h += record_hits(projdata + 20u * (v * mods + k), lo, hi, vox, n, double(a->line_offset));
            worst = std::fmax(worst, h);
        }
    }
    *hits = worst;
    return XPR_OK;
}

extern "C" int xpr_geometry(const xpr_geometry_params *p, float *projdata_out,
                            xpr_grid *grid_out)
{
    if (!p || !projdata_out || !grid_out) return xprint_fail(XPR_E_INVALID, "null pointer");
    if (p->module_number < 1 || p->dwidth < 1 || p->dheight < 1 || p->projections < 1)
        return xprint_fail(XPR_E_INVALID, "geometry: empty detector or no projection");
    if (!(p->so > 0.0) || !(p->sd > 0.0) || !(p->px > 0.0))
        return xprint_fail(XPR_E_INVALID, "geometry: so, sd and px must be positive");
    if (p->fp64)
        xpr_geom::build<double>(*p, projdata_out, *grid_out);
    else
        xpr_geom::build<float>(*p, projdata_out, *grid_out);
    return XPR_OK;
}

extern "C" int xpr_geometry_frames(const xpr_geometry_params *p, double *frames_out)
{
    if (!p || !frames_out) return xprint_fail(XPR_E_INVALID, "null pointer");
    if (p->module_number < 1 || p->dwidth < 1 || p->dheight < 1 || p->projections < 1)
        return xprint_fail(XPR_E_INVALID, "geometry: empty detector or no projection");
    xpr_grid g;
    xpr_geom::build<double>(*p, nullptr, g, frames_out);
    return XPR_OK;
}
