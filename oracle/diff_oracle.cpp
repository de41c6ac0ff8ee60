// ORACLE — test infrastructure only.
//
// FP64 CPU reference of the differentiable render: forward, analytic
// gradient of an L2 image loss, and finite differences of the same loss.
// splat-tui has no backward; the forward semantics it differentiates are the
// `cpu-naive` rules (src/render/naive.rs:29-100, src/render/mod.rs:201-285):
// culling (opacity < ALPHA_MIN, depth <= near, det <= 0, off-screen,
// non-finite), pixel boxes of k*sqrt(a), k*sqrt(c) (src/render/cpu.rs:96-102),
// per pixel in (depth, index) order: skip power > 0 and alpha < 1/255,
// alpha = min(opacity * exp(power), 0.99), stop after T < 1e-4. Parameters
// are raw: means, log-scales, quaternions (normalised), opacity logits, base
// colours, SH rest coefficients (src/scene.rs:62-94 activates them the same
// way: exp, sigmoid, normalise).
//
// The analytic gradient is written in matrix form (Sigma' = T Sigma T^T with
// T = J W, conic = Sigma'^-1, Sigma = (R S)(R S)^T), not in the row-vector
// form of the HIP adjoint, so the two are independent derivations; the
// finite differences (or_diff_fd) pin both.
//
// Only the tests and the benchmark's gradient check load this library.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "splat_hip.h"

namespace {

constexpr double AMIN = (double)(1.0f / 255.0f);
constexpr double AMAX = (double)0.99f;
constexpr double TMIN = (double)1e-4f;
constexpr double DIL = (double)0.3f;
constexpr uint32_t TILE = 16;

template <class F>
void par_for(size_t n, int threads, F &&fn) {
    if (threads <= 1 || n <= 1) {
        for (size_t i = 0; i < n; ++i) fn(i, 0);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i, t);
        });
    for (auto &th : pool) th.join();
}

struct View64 {
    double W[3][3];  // rows right, down, forward
    double eye[3];
    double focal, cx, cy, limx, limy, near_z;
    uint32_t w, h, tiles_x, tiles_y;
};

View64 make_view(const sp_view &v, uint32_t w, uint32_t h) {
    View64 o;
    for (int c = 0; c < 3; ++c) {
        o.W[0][c] = v.right[c];
        o.W[1][c] = v.down[c];
        o.W[2][c] = v.forward[c];
        o.eye[c] = v.eye[c];
    }
    o.focal = 0.5 * h / std::tan(0.5 * (double)v.fov_y);
    o.cx = 0.5 * w;
    o.cy = 0.5 * h;
    o.limx = 1.3 * (0.5 * w / o.focal);
    o.limy = 1.3 * (0.5 * h / o.focal);
    if (!std::isfinite(o.limx)) o.limx = 0;
    if (!std::isfinite(o.limy)) o.limy = 0;
    o.near_z = v.near_z;
    o.w = w;
    o.h = h;
    o.tiles_x = (w + TILE - 1) / TILE;
    o.tiles_y = (h + TILE - 1) / TILE;
    return o;
}

// Real SH basis (degree <= 3, the 15 non-DC terms of src/sh.rs:39-64) and its
// gradient in the direction components.
void sh_eval(uint32_t deg, const double d[3], double Y[15], double dY[15][3]) {
    const double x = d[0], y = d[1], z = d[2];
    const double C1 = 0.4886025119029199;
    const double C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
                          0.5462742152960396};
    const double C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                          -0.4570457994644658, 1.445305721320277, -0.5900435899266435};
    auto set = [&](int k, double v, double gx, double gy, double gz) {
        Y[k] = v;
        dY[k][0] = gx;
        dY[k][1] = gy;
        dY[k][2] = gz;
    };
    if (deg >= 1) {
        set(0, -C1 * y, 0, -C1, 0);
        set(1, C1 * z, 0, 0, C1);
        set(2, -C1 * x, -C1, 0, 0);
    }
    if (deg >= 2) {
        set(3, C2[0] * x * y, C2[0] * y, C2[0] * x, 0);
        set(4, C2[1] * y * z, 0, C2[1] * z, C2[1] * y);
        set(5, C2[2] * (2 * z * z - x * x - y * y), -2 * C2[2] * x, -2 * C2[2] * y, 4 * C2[2] * z);
        set(6, C2[3] * x * z, C2[3] * z, 0, C2[3] * x);
        set(7, C2[4] * (x * x - y * y), 2 * C2[4] * x, -2 * C2[4] * y, 0);
    }
    if (deg >= 3) {
        // Products expanded to monomials before differentiating.
        set(8, C3[0] * (3 * x * x * y - y * y * y), C3[0] * 6 * x * y, C3[0] * (3 * x * x - 3 * y * y), 0);
        set(9, C3[1] * x * y * z, C3[1] * y * z, C3[1] * x * z, C3[1] * x * y);
        set(10, C3[2] * (4 * y * z * z - x * x * y - y * y * y), C3[2] * (-2 * x * y),
            C3[2] * (4 * z * z - x * x - 3 * y * y), C3[2] * 8 * y * z);
        set(11, C3[3] * (2 * z * z * z - 3 * x * x * z - 3 * y * y * z), C3[3] * (-6 * x * z), C3[3] * (-6 * y * z),
            C3[3] * (6 * z * z - 3 * x * x - 3 * y * y));
        set(12, C3[4] * (4 * x * z * z - x * x * x - x * y * y), C3[4] * (4 * z * z - 3 * x * x - y * y),
            C3[4] * (-2 * x * y), C3[4] * 8 * x * z);
        set(13, C3[5] * (x * x * z - y * y * z), C3[5] * 2 * x * z, C3[5] * (-2 * y * z), C3[5] * (x * x - y * y));
        set(14, C3[6] * (x * x * x - 3 * x * y * y), C3[6] * (3 * x * x - 3 * y * y), C3[6] * (-6 * x * y), 0);
    }
}

// Rotation of a unit quaternion (x, y, z, w) and its derivatives.
void quat_rot(const double q[4], double R[3][3]) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double M[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                            {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                            {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
    std::memcpy(R, M, sizeof M);
}
void quat_rot_grad(const double q[4], double D[4][3][3]) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double Dx[3][3] = {{0, 2 * y, 2 * z}, {2 * y, -4 * x, -2 * w}, {2 * z, 2 * w, -4 * x}};
    const double Dy[3][3] = {{-4 * y, 2 * x, 2 * w}, {2 * x, 0, 2 * z}, {-2 * w, 2 * z, -4 * y}};
    const double Dz[3][3] = {{-4 * z, -2 * w, 2 * x}, {2 * w, -4 * z, 2 * y}, {2 * x, 2 * y, 0}};
    const double Dw[3][3] = {{0, -2 * z, 2 * y}, {2 * z, 0, -2 * x}, {-2 * y, 2 * x, 0}};
    std::memcpy(D[0], Dx, sizeof Dx);
    std::memcpy(D[1], Dy, sizeof Dy);
    std::memcpy(D[2], Dz, sizeof Dz);
    std::memcpy(D[3], Dw, sizeof Dw);
}

// One splat projected into one view, with what the backward needs.
struct Proj {
    bool ok = false;
    double x, y, A, B, C, op, col[3], pre[3], depth;
    uint32_t bb[4];
    // intermediates
    double t[3], f, jx, jy, T[2][3], Sig[3][3], a, b, c, det, dir[3], len, s[3], qn[4], qlen, R[3][3];
    bool freex, freey, qok, dirok;
};

struct Layout {
    uint32_t deg, rest, np;  // np = 14 + 3 * rest parameters per splat
};

// p: the splat's np raw parameters [mean 3, log-scale 3, quat 4, logit, colour 3, sh 3*rest].
void project(const Layout &L, const double *p, const View64 &v, Proj &o) {
    o.ok = false;
    for (int c = 0; c < 3; ++c) o.s[c] = std::exp(p[3 + c]);
    o.qlen = std::sqrt(p[6] * p[6] + p[7] * p[7] + p[8] * p[8] + p[9] * p[9]);
    o.qok = std::isfinite(o.qlen) && o.qlen > 1e-12;
    for (int c = 0; c < 4; ++c) o.qn[c] = o.qok ? p[6 + c] / o.qlen : (c == 3 ? 1.0 : 0.0);
    quat_rot(o.qn, o.R);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += o.R[r][k] * o.s[k] * o.s[k] * o.R[c][k];
            o.Sig[r][c] = acc;
        }
    o.op = 1.0 / (1.0 + std::exp(-p[10]));
    if (o.op < AMIN) return;
    double d[3];
    for (int c = 0; c < 3; ++c) d[c] = p[c] - v.eye[c];
    for (int r = 0; r < 3; ++r) o.t[r] = v.W[r][0] * d[0] + v.W[r][1] * d[1] + v.W[r][2] * d[2];
    if (!(o.t[2] > v.near_z)) return;
    o.x = v.focal * o.t[0] / o.t[2] + v.cx;
    o.y = v.focal * o.t[1] / o.t[2] + v.cy;
    const double jxr = o.t[0] / o.t[2], jyr = o.t[1] / o.t[2];
    o.freex = jxr >= -v.limx && jxr <= v.limx;
    o.freey = jyr >= -v.limy && jyr <= v.limy;
    o.jx = std::min(std::max(jxr, -v.limx), v.limx);
    o.jy = std::min(std::max(jyr, -v.limy), v.limy);
    o.f = v.focal / o.t[2];
    const double J[2][3] = {{o.f, 0, -o.f * o.jx}, {0, o.f, -o.f * o.jy}};
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) o.T[r][c] = J[r][0] * v.W[0][c] + J[r][1] * v.W[1][c] + J[r][2] * v.W[2][c];
    double S2[2][2];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
            double acc = 0;
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) acc += o.T[r][k] * o.Sig[k][l] * o.T[c][l];
            S2[r][c] = acc;
        }
    o.a = S2[0][0] + DIL;
    o.b = S2[0][1];
    o.c = S2[1][1] + DIL;
    o.det = o.a * o.c - o.b * o.b;
    if (!(o.det > 0)) return;
    const double k = std::sqrt(2.0 * std::log(o.op / AMIN));
    const double ex = k * std::sqrt(o.a), ey = k * std::sqrt(o.c);
    if (o.x + ex <= 0 || o.x - ex >= v.w || o.y + ey <= 0 || o.y - ey >= v.h) return;
    if (!std::isfinite(o.x) || !std::isfinite(o.y) || !std::isfinite(ex) || !std::isfinite(ey)) return;
    for (int c = 0; c < 3; ++c) o.pre[c] = p[11 + c];
    o.len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    o.dirok = L.deg > 0 && std::isfinite(1.0 / o.len) && o.len > 0;
    for (int c = 0; c < 3; ++c) o.dir[c] = o.dirok ? d[c] / o.len : 0.0;
    if (L.deg > 0) {
        double Y[15], dY[15][3];
        sh_eval(L.deg, o.dir, Y, dY);
        for (uint32_t q = 0; q < L.rest; ++q)
            for (int ch = 0; ch < 3; ++ch) o.pre[ch] += Y[q] * p[14 + 3 * q + ch];
    }
    for (int c = 0; c < 3; ++c) o.col[c] = std::max(o.pre[c], 0.0);
    o.A = o.c / o.det;
    o.B = -o.b / o.det;
    o.C = o.a / o.det;
    o.depth = o.t[2];
    const double fx0 = std::max(std::floor(o.x - ex), 0.0), fy0 = std::max(std::floor(o.y - ey), 0.0);
    const uint32_t x1 = (uint32_t)std::min(std::max(std::ceil(o.x + ex), 0.0), (double)v.w);
    const uint32_t y1 = (uint32_t)std::min(std::max(std::ceil(o.y + ey), 0.0), (double)v.h);
    o.bb[0] = std::min((uint32_t)std::min(fx0, 4.0e9), x1);
    o.bb[1] = std::min((uint32_t)std::min(fy0, 4.0e9), y1);
    o.bb[2] = x1;
    o.bb[3] = y1;
    o.ok = true;
}

// Per-pixel blend helpers shared by the forward, backward and FD paths.
inline bool alpha_of(const Proj &p, double px, double py, double *al, double *ex, double *dx, double *dy) {
    *dx = px + 0.5 - p.x;
    *dy = py + 0.5 - p.y;
    const double pw = -0.5 * (p.A * *dx * *dx + p.C * *dy * *dy) - p.B * *dx * *dy;
    if (pw > 0) return false;
    *ex = std::exp(pw);
    *al = std::min(p.op * *ex, AMAX);
    return *al >= AMIN;
}

// 2D gradient of one record: x, y, A, B, C, opacity, r, g, b.
constexpr int NG = 9;

// Chain the record's 2D gradient through the projection; adds to gp (np).
void project_backward(const Layout &L, const double *p, const View64 &v, const Proj &o, const double g[NG],
                      double *gp) {
    // conic = Sigma'^-1: dL/dSigma' = -Sigma'^-1 G_inv Sigma'^-1, G_inv = [[gA, gB/2], [gB/2, gC]].
    const double Si[2][2] = {{o.A, o.B}, {o.B, o.C}};
    const double Gi[2][2] = {{g[2], 0.5 * g[3]}, {0.5 * g[3], g[4]}};
    double tmp[2][2], Gp[2][2];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) tmp[r][c] = Gi[r][0] * Si[0][c] + Gi[r][1] * Si[1][c];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) Gp[r][c] = -(Si[r][0] * tmp[0][c] + Si[r][1] * tmp[1][c]);
    // Sigma' = T Sigma T^T (+ dilation): dL/dT = (Gp + Gp^T) T Sigma, dL/dSigma = T^T Gp T.
    double GT[2][3], gSig[3][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
            for (int k = 0; k < 2; ++k)
                for (int l = 0; l < 3; ++l) acc += (Gp[r][k] + Gp[k][r]) * o.T[k][l] * o.Sig[l][c];
            GT[r][c] = acc;
        }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
            for (int k = 0; k < 2; ++k)
                for (int l = 0; l < 2; ++l) acc += o.T[k][r] * Gp[k][l] * o.T[l][c];
            gSig[r][c] = acc;
        }
    // T = J W: dL/dJ = dL/dT W^T.
    double GJ[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) GJ[r][c] = GT[r][0] * v.W[c][0] + GT[r][1] * v.W[c][1] + GT[r][2] * v.W[c][2];
    // J = [[f, 0, -f jx], [0, f, -f jy]], f = focal / tz, jx = tx / tz (unless clamped).
    const double gf = GJ[0][0] + GJ[1][1] - GJ[0][2] * o.jx - GJ[1][2] * o.jy;
    const double gjx = -o.f * GJ[0][2], gjy = -o.f * GJ[1][2];
    const double tz = o.t[2];
    double gt[3] = {0, 0, 0};
    gt[2] += -gf * v.focal / (tz * tz);
    if (o.freex) {
        gt[0] += gjx / tz;
        gt[2] += -gjx * o.t[0] / (tz * tz);
    }
    if (o.freey) {
        gt[1] += gjy / tz;
        gt[2] += -gjy * o.t[1] / (tz * tz);
    }
    // Mean: x = focal tx / tz + cx, y = focal ty / tz + cy.
    gt[0] += g[0] * v.focal / tz;
    gt[1] += g[1] * v.focal / tz;
    gt[2] += -(g[0] * v.focal * o.t[0] + g[1] * v.focal * o.t[1]) / (tz * tz);
    double gd[3];
    for (int c = 0; c < 3; ++c) gd[c] = v.W[0][c] * gt[0] + v.W[1][c] * gt[1] + v.W[2][c] * gt[2];
    // Colour: col = max(base + sum_k Y_k(dir) sh_k, 0).
    double gc[3];
    for (int c = 0; c < 3; ++c) gc[c] = o.pre[c] > 0 ? g[6 + c] : 0.0;
    for (int c = 0; c < 3; ++c) gp[11 + c] += gc[c];
    if (L.deg > 0) {
        double Y[15], dY[15][3], gdir[3] = {0, 0, 0};
        sh_eval(L.deg, o.dir, Y, dY);
        for (uint32_t q = 0; q < L.rest; ++q)
            for (int ch = 0; ch < 3; ++ch) {
                gp[14 + 3 * q + ch] += gc[ch] * Y[q];
                for (int c = 0; c < 3; ++c) gdir[c] += gc[ch] * p[14 + 3 * q + ch] * dY[q][c];
            }
        if (o.dirok) {
            // dir = d / |d|: d dir / d d = (I - dir dir^T) / |d|.
            for (int r = 0; r < 3; ++r) {
                double acc = 0;
                for (int c = 0; c < 3; ++c) acc += ((r == c ? 1.0 : 0.0) - o.dir[r] * o.dir[c]) * gdir[c];
                gd[r] += acc / o.len;
            }
        }
    }
    for (int c = 0; c < 3; ++c) gp[c] += gd[c];
    // Opacity logit.
    gp[10] += g[5] * o.op * (1 - o.op);
    // Sigma = M M^T, M = R diag(s): dL/dM = (gSig + gSig^T) M.
    double M[3][3], gM[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] = o.R[r][c] * o.s[c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += (gSig[r][k] + gSig[k][r]) * M[k][c];
            gM[r][c] = acc;
        }
    double gR[3][3];
    for (int c = 0; c < 3; ++c) {
        double gs = 0;
        for (int r = 0; r < 3; ++r) {
            gs += gM[r][c] * o.R[r][c];
            gR[r][c] = gM[r][c] * o.s[c];
        }
        gp[3 + c] += gs * o.s[c];  // log-scale
    }
    if (o.qok) {
        double D[4][3][3], gqn[4];
        quat_rot_grad(o.qn, D);
        for (int k = 0; k < 4; ++k) {
            double acc = 0;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) acc += gR[r][c] * D[k][r][c];
            gqn[k] = acc;
        }
        double dot = 0;
        for (int k = 0; k < 4; ++k) dot += o.qn[k] * gqn[k];
        for (int k = 0; k < 4; ++k) gp[6 + k] += (gqn[k] - o.qn[k] * dot) / o.qlen;
    }
}

}  // namespace

struct or_diff {
    Layout L;
    size_t n = 0;
    std::vector<double> P;  // n x np raw parameters
    double bg[3];
    int threads = 1;
    struct PerView {
        View64 v;
        std::vector<Proj> pr;               // n
        std::vector<uint32_t> tile_start;   // tiles + 1
        std::vector<uint32_t> list;         // scene indices, per tile in (depth, index) order
        std::vector<double> img;            // h * w * 3
        std::vector<double> Tfin;           // h * w
        std::vector<uint32_t> last;         // h * w: list position after the last contribution
        std::vector<float> target;          // h * w * 3 (empty: zero target)
    };
    std::vector<PerView> views;
    double loss = 0;
};

namespace {

double target_of(const or_diff::PerView &V, size_t q, int c) {
    return V.target.empty() ? 0.0 : (double)V.target[3 * q + c];
}

void build_view(or_diff *D, or_diff::PerView &V) {
    const size_t n = D->n;
    V.pr.assign(n, Proj());
    par_for((n + 4095) / 4096, D->threads, [&](size_t c, int) {
        for (size_t i = c * 4096; i < std::min(n, (c + 1) * 4096); ++i)
            project(D->L, &D->P[i * D->L.np], V.v, V.pr[i]);
    });
    std::vector<uint32_t> order;
    for (size_t i = 0; i < n; ++i)
        if (V.pr[i].ok) order.push_back((uint32_t)i);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return V.pr[a].depth < V.pr[b].depth; });
    const uint32_t tiles = V.v.tiles_x * V.v.tiles_y;
    std::vector<uint32_t> cnt(tiles + 1, 0);
    auto rect = [&](const Proj &p, uint32_t &tx0, uint32_t &ty0, uint32_t &tx1, uint32_t &ty1) {
        tx0 = p.bb[0] / TILE;
        ty0 = p.bb[1] / TILE;
        tx1 = (p.bb[2] + TILE - 1) / TILE;
        ty1 = (p.bb[3] + TILE - 1) / TILE;
        if (p.bb[0] >= p.bb[2] || p.bb[1] >= p.bb[3]) tx1 = tx0, ty1 = ty0;
    };
    for (uint32_t i : order) {
        uint32_t tx0, ty0, tx1, ty1;
        rect(V.pr[i], tx0, ty0, tx1, ty1);
        for (uint32_t ty = ty0; ty < ty1; ++ty)
            for (uint32_t tx = tx0; tx < tx1; ++tx) cnt[ty * V.v.tiles_x + tx + 1]++;
    }
    for (uint32_t t = 0; t < tiles; ++t) cnt[t + 1] += cnt[t];
    V.tile_start = cnt;
    V.list.assign(cnt[tiles], 0);
    std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
    for (uint32_t i : order) {
        uint32_t tx0, ty0, tx1, ty1;
        rect(V.pr[i], tx0, ty0, tx1, ty1);
        for (uint32_t ty = ty0; ty < ty1; ++ty)
            for (uint32_t tx = tx0; tx < tx1; ++tx) V.list[fill[ty * V.v.tiles_x + tx]++] = i;
    }
    const size_t npx = (size_t)V.v.w * V.v.h;
    V.img.assign(3 * npx, 0.0);
    V.Tfin.assign(npx, 1.0);
    V.last.assign(npx, 0);
    par_for(tiles, D->threads, [&](size_t t, int) {
        const uint32_t tx = (uint32_t)t % V.v.tiles_x, ty = (uint32_t)t / V.v.tiles_x;
        const uint32_t s0 = V.tile_start[t], s1 = V.tile_start[t + 1];
        for (uint32_t py = ty * TILE; py < std::min(V.v.h, ty * TILE + TILE); ++py)
            for (uint32_t px = tx * TILE; px < std::min(V.v.w, tx * TILE + TILE); ++px) {
                double T = 1, C[3] = {0, 0, 0};
                uint32_t last = 0;
                for (uint32_t k = s0; k < s1 && T >= TMIN; ++k) {
                    const Proj &p = V.pr[V.list[k]];
                    if (px < p.bb[0] || px >= p.bb[2] || py < p.bb[1] || py >= p.bb[3]) continue;
                    double al, ex, dx, dy;
                    if (!alpha_of(p, px, py, &al, &ex, &dx, &dy)) continue;
                    for (int c = 0; c < 3; ++c) C[c] += p.col[c] * al * T;
                    T *= 1 - al;
                    last = k - s0 + 1;
                }
                const size_t q = (size_t)py * V.v.w + px;
                for (int c = 0; c < 3; ++c) V.img[3 * q + c] = C[c] + T * D->bg[c];
                V.Tfin[q] = T;
                V.last[q] = last;
            }
    });
}

double view_loss(const or_diff::PerView &V) {
    double l = 0;
    for (size_t q = 0; q < (size_t)V.v.w * V.v.h; ++q)
        for (int c = 0; c < 3; ++c) {
            const double r = V.img[3 * q + c] - target_of(V, q, c);
            l += r * r;
        }
    return l;
}

// Loss over the pixels of rectangle [x0,x1)x[y0,y1) of view V with splat `si`
// replaced by `ps` (ps.ok == false: removed).
double region_loss(const or_diff *D, const or_diff::PerView &V, uint32_t si, const Proj &ps, uint32_t x0,
                   uint32_t y0, uint32_t x1, uint32_t y1) {
    double l = 0;
    for (uint32_t py = y0; py < y1; ++py)
        for (uint32_t px = x0; px < x1; ++px) {
            const uint32_t t = (py / TILE) * V.v.tiles_x + px / TILE;
            const uint32_t s0 = V.tile_start[t], s1 = V.tile_start[t + 1];
            double T = 1, C[3] = {0, 0, 0};
            bool placed = !ps.ok;
            auto blend = [&](const Proj &p) {
                if (px < p.bb[0] || px >= p.bb[2] || py < p.bb[1] || py >= p.bb[3]) return;
                double al, ex, dx, dy;
                if (!alpha_of(p, px, py, &al, &ex, &dx, &dy)) return;
                for (int c = 0; c < 3; ++c) C[c] += p.col[c] * al * T;
                T *= 1 - al;
            };
            for (uint32_t k = s0; k < s1 && T >= TMIN; ++k) {
                const uint32_t j = V.list[k];
                if (j == si) continue;
                const Proj &p = V.pr[j];
                if (!placed && (ps.depth < p.depth || (ps.depth == p.depth && si < j))) {
                    placed = true;
                    blend(ps);
                    if (T < TMIN) break;
                }
                blend(p);
            }
            if (!placed && T >= TMIN) blend(ps);
            const size_t q = (size_t)py * V.v.w + px;
            for (int c = 0; c < 3; ++c) {
                const double r = C[c] + T * D->bg[c] - target_of(V, q, c);
                l += r * r;
            }
        }
    return l;
}

}  // namespace

extern "C" int or_resolve_threads(int threads);  // splat_oracle.cpp

extern "C" {

// Build the FP64 session: params as float arrays in the sp_params layout
// (host memory), views with sizes wh, optional target images (float, h*w*3
// per view; NULL entries mean zero targets). Runs the forward and the loss
// L = sum over views, pixels, channels of (image - target)^2.
or_diff *or_diff_create(const float *means, const float *log_scales, const float *quats, const float *logits,
                        const float *colors, const float *sh_rest, uint64_t n, uint32_t deg, const sp_view *views,
                        const uint32_t *wh, uint32_t nviews, const float bg[3], const float *const *targets,
                        int threads) {
    or_diff *D = new or_diff();
    D->L.deg = deg;
    D->L.rest = (deg + 1) * (deg + 1) - 1;
    D->L.np = 14 + 3 * D->L.rest;
    D->n = n;
    D->threads = or_resolve_threads(threads);
    for (int c = 0; c < 3; ++c) D->bg[c] = bg[c];
    const uint32_t np = D->L.np;
    D->P.assign(n * np, 0.0);
    for (size_t i = 0; i < n; ++i) {
        double *p = &D->P[i * np];
        for (int c = 0; c < 3; ++c) {
            p[c] = means[3 * i + c];
            p[3 + c] = log_scales[3 * i + c];
            p[11 + c] = colors[3 * i + c];
        }
        for (int c = 0; c < 4; ++c) p[6 + c] = quats[4 * i + c];
        p[10] = logits[i];
        for (uint32_t k = 0; k < 3 * D->L.rest; ++k) p[14 + k] = sh_rest[3 * D->L.rest * i + k];
    }
    D->views.resize(nviews);
    D->loss = 0;
    for (uint32_t v = 0; v < nviews; ++v) {
        auto &V = D->views[v];
        V.v = make_view(views[v], wh[2 * v], wh[2 * v + 1]);
        if (targets && targets[v]) V.target.assign(targets[v], targets[v] + 3ull * wh[2 * v] * wh[2 * v + 1]);
        build_view(D, V);
        D->loss += view_loss(V);
    }
    return D;
}

void or_diff_destroy(or_diff *D) { delete D; }

double or_diff_loss(const or_diff *D) { return D->loss; }

// The FP64 forward image of view v (h*w*3 doubles).
void or_diff_image(const or_diff *D, uint32_t v, double *out) {
    std::copy(D->views[v].img.begin(), D->views[v].img.end(), out);
}

// Analytic dL/dparameters, n x np doubles in the per-splat order
// [mean 3, log-scale 3, quat 4, logit, colour 3, sh 3*rest].
void or_diff_grad(const or_diff *D, double *grad) {
    const uint32_t np = D->L.np;
    const size_t n = D->n;
    std::fill(grad, grad + n * np, 0.0);
    const int nt = D->threads;
    for (const auto &V : D->views) {
        const uint32_t tiles = V.v.tiles_x * V.v.tiles_y;
        // Per-thread 2D accumulators over contiguous tile chunks, summed in a fixed order.
        const size_t chunks = (size_t)nt;
        std::vector<std::vector<double>> acc(chunks, std::vector<double>(n * NG, 0.0));
        par_for(chunks, nt, [&](size_t ch, int) {
            std::vector<double> &g2 = acc[ch];
            const uint32_t t0 = (uint32_t)(tiles * ch / chunks), t1 = (uint32_t)(tiles * (ch + 1) / chunks);
            for (uint32_t t = t0; t < t1; ++t) {
                const uint32_t tx = t % V.v.tiles_x, ty = t / V.v.tiles_x;
                const uint32_t s0 = V.tile_start[t];
                for (uint32_t py = ty * TILE; py < std::min(V.v.h, ty * TILE + TILE); ++py)
                    for (uint32_t px = tx * TILE; px < std::min(V.v.w, tx * TILE + TILE); ++px) {
                        const size_t q = (size_t)py * V.v.w + px;
                        double G[3], S[3], T = V.Tfin[q];
                        for (int c = 0; c < 3; ++c) {
                            G[c] = 2 * (V.img[3 * q + c] - target_of(V, q, c));
                            S[c] = T * D->bg[c];
                        }
                        for (uint32_t k = V.last[q]; k-- > 0;) {
                            const uint32_t j = V.list[s0 + k];
                            const Proj &p = V.pr[j];
                            if (px < p.bb[0] || px >= p.bb[2] || py < p.bb[1] || py >= p.bb[3]) continue;
                            double al, ex, dx, dy;
                            if (!alpha_of(p, px, py, &al, &ex, &dx, &dy)) continue;
                            const double Ti = T / (1 - al);
                            double ga = 0;
                            for (int c = 0; c < 3; ++c) ga += G[c] * (p.col[c] * Ti - S[c] / (1 - al));
                            double *o = &g2[(size_t)j * NG];
                            for (int c = 0; c < 3; ++c) {
                                o[6 + c] += G[c] * al * Ti;
                                S[c] += p.col[c] * al * Ti;
                            }
                            T = Ti;
                            if (p.op * ex < AMAX) {
                                o[5] += ga * ex;
                                const double gpw = ga * al;
                                o[0] += gpw * (p.A * dx + p.B * dy);
                                o[1] += gpw * (p.B * dx + p.C * dy);
                                o[2] += gpw * -0.5 * dx * dx;
                                o[3] += gpw * -dx * dy;
                                o[4] += gpw * -0.5 * dy * dy;
                            }
                        }
                    }
            }
        });
        par_for((n + 4095) / 4096, nt, [&](size_t c, int) {
            for (size_t i = c * 4096; i < std::min(n, (c + 1) * 4096); ++i) {
                if (!V.pr[i].ok) continue;
                double g[NG] = {0};
                bool any = false;
                for (size_t ch = 0; ch < chunks; ++ch)
                    for (int k = 0; k < NG; ++k) g[k] += acc[ch][i * NG + k];
                for (int k = 0; k < NG; ++k) any = any || g[k] != 0;
                if (any) project_backward(D->L, &D->P[i * np], V.v, V.pr[i], g, grad + i * np);
            }
        });
    }
}

// Central finite difference of L with respect to parameter k of splat i,
// (L(p + eps e_k) - L(p - eps e_k)) / (2 eps), evaluated over the pixels the
// splat can touch before or after the step (the rest of the image does not
// change).
double or_diff_fd(const or_diff *D, uint64_t i, uint32_t k, double eps) {
    const uint32_t np = D->L.np;
    std::vector<double> pp(D->P.begin() + i * np, D->P.begin() + (i + 1) * np), pm = pp;
    pp[k] += eps;
    pm[k] -= eps;
    double lp = 0, lm = 0;
    for (const auto &V : D->views) {
        Proj a, b;
        project(D->L, pp.data(), V.v, a);
        project(D->L, pm.data(), V.v, b);
        const Proj &o = V.pr[i];
        uint32_t x0 = V.v.w, y0 = V.v.h, x1 = 0, y1 = 0;
        for (const Proj *q : {&o, (const Proj *)&a, (const Proj *)&b}) {
            if (!q->ok) continue;
            x0 = std::min(x0, q->bb[0]);
            y0 = std::min(y0, q->bb[1]);
            x1 = std::max(x1, q->bb[2]);
            y1 = std::max(y1, q->bb[3]);
        }
        if (x0 >= x1 || y0 >= y1) continue;
        lp += region_loss(D, V, (uint32_t)i, a, x0, y0, x1, y1);
        lm += region_loss(D, V, (uint32_t)i, b, x0, y0, x1, y1);
    }
    return (lp - lm) / (2 * eps);
}

uint32_t or_diff_nparams(const or_diff *D) { return D->L.np; }

}  // extern "C"
