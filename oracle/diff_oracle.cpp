// diff_oracle.cpp -- CPU oracle of the differentiable renderer (TEST INFRASTRUCTURE ONLY).
//
// The reference has no differentiable mode; this file states the model the GPU adjoint
// kernels implement, in plain loops, templated on the scalar type: the float build
// follows the device arithmetic step for step, the double build is the accurate model
// that finite differences are taken on.
//
// Model.  Paths of line / quadratic / cubic segments (closed contours), fill rule
// nonzero or even-odd, paint RGBA in [0,1] (linear).  A pixel row y holds S sample rows
// at y + (s + 1/2)/S; along each sample row the fill indicator is integrated exactly over
// the pixel's [x, x+1) (a box filter: S rows, continuous in x), so
//     cov_p(x, y) = (1/S) * sum_s |{u in [x, x+1) : fill_p(u, y_s)}|.
// Compositing is the renderer's src-over onto white in path order, without quantisation:
//     a = paint_a * cov;  rgb = paint_rgb * a + rgb * (1 - a);  A = a * a + A * (1 - a).
// Crossings of sample row y_s with a segment are found on its y-monotone pieces under the
// half-open rule ymin <= y_s < ymax, by bisection on the piece's t range.
//
// Gradients.  Colours: the reverse of the compositing chain.  Control points: boundary
// terms at the crossings -- d cov / d x_i = (fill left of x_i - fill right of x_i) / S in
// the pixel holding x_i, and x_i = Bx(t*) with By(t*) = y_s gives
//     d x_i / d Pk.x = b_k(t*),   d x_i / d Pk.y = -Bx'(t*) / By'(t*) * b_k(t*)
// with b_k the Bernstein weights of the segment.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <class T> struct Seg {
    int type;            // 2 line, 3 quadratic, 4 cubic
    T x[4], y[4];
};

template <class T> T bez(int type, const T *p, T t) {
    const T s = T(1) - t;
    if (type == 2) return s * p[0] + t * p[1];
    if (type == 3) return s * s * p[0] + T(2) * s * t * p[1] + t * t * p[2];
    return s * s * s * p[0] + T(3) * s * s * t * p[1] + T(3) * s * t * t * p[2] + t * t * t * p[3];
}

template <class T> T dbez(int type, const T *p, T t) {
    const T s = T(1) - t;
    if (type == 2) return p[1] - p[0];
    if (type == 3) return T(2) * (s * (p[1] - p[0]) + t * (p[2] - p[1]));
    return T(3) * (s * s * (p[1] - p[0]) + T(2) * s * t * (p[2] - p[1]) + t * t * (p[3] - p[2]));
}

template <class T> void basis(int type, T t, T *b) {
    const T s = T(1) - t;
    b[0] = b[1] = b[2] = b[3] = T(0);
    if (type == 2) { b[0] = s; b[1] = t; return; }
    if (type == 3) { b[0] = s * s; b[1] = T(2) * s * t; b[2] = t * t; return; }
    b[0] = s * s * s; b[1] = T(3) * s * s * t; b[2] = T(3) * s * t * t; b[3] = t * t * t;
}

// t of the y-extrema inside (0,1), sorted; returns the count (<= 2)
template <class T> int y_extrema(int type, const T *y, T *r) {
    int n = 0;
    if (type == 3) {
        const T den = y[0] - T(2) * y[1] + y[2];
        if (den != T(0)) {
            const T t = (y[0] - y[1]) / den;
            if (t > T(0) && t < T(1)) r[n++] = t;
        }
    } else if (type == 4) {
        const T a = -y[0] + T(3) * y[1] - T(3) * y[2] + y[3];
        const T b = T(2) * (y[0] - T(2) * y[1] + y[2]);
        const T c = y[1] - y[0];
        if (a == T(0)) {
            if (b != T(0)) {
                const T t = -c / b;
                if (t > T(0) && t < T(1)) r[n++] = t;
            }
        } else {
            const T disc = b * b - T(4) * a * c;
            if (disc > T(0)) {
                const T sq = std::sqrt(disc);
                T t0 = (-b - sq) / (T(2) * a), t1 = (-b + sq) / (T(2) * a);
                if (t0 > t1) std::swap(t0, t1);
                if (t0 > T(0) && t0 < T(1)) r[n++] = t0;
                if (t1 > T(0) && t1 < T(1) && t1 != t0) r[n++] = t1;
            }
        }
    }
    return n;
}

template <class T> struct Cross {
    T x, t;
    int curve, row, dir;  // sample row, +1 downward (increasing y), -1 upward
    int path;
};

// first sample row index j with (j + 1/2)/S >= y, clamped to [-1, J + 1] (NaN -> -1)
template <class T> long first_row(T y, int S, long J) {
    T v = std::ceil(y * T(S) - T(0.5));
    if (!(v >= T(-1))) v = T(-1);
    if (v > T(J + 1)) v = T(J + 1);
    return (long)v;
}

template <class T>
void crossings(const std::vector<Seg<T>> &segs, const std::vector<int> &cpath, int H, int S,
               std::vector<Cross<T>> &out) {
    const long J = (long)H * S;
    for (size_t c = 0; c < segs.size(); ++c) {
        const Seg<T> &g = segs[c];
        T ts[4] = {T(0), T(0), T(0), T(0)};
        const int ne = y_extrema(g.type, g.y, ts + 1);
        ts[ne + 1] = T(1);
        T yprev = bez(g.type, g.y, T(0));
        for (int k = 0; k <= ne; ++k) {
            const T t0 = ts[k], t1 = ts[k + 1];
            const T y0 = yprev, y1 = bez(g.type, g.y, t1);
            yprev = y1;
            if (y0 == y1) continue;
            const T ylo = std::min(y0, y1), yhi = std::max(y0, y1);
            long ja = std::max(first_row(ylo, S, J), 0L), jb = std::min(first_row(yhi, S, J), J);
            const int dir = y1 > y0 ? 1 : -1;
            for (long j = ja; j < jb; ++j) {
                const T ys = (T(j) + T(0.5)) / T(S);
                // bisection with By(lo) < ys <= By(hi) in the piece's increasing direction
                T lo = dir > 0 ? t0 : t1, hi = dir > 0 ? t1 : t0;
                const int iters = sizeof(T) == 4 ? 32 : 64;
                for (int it = 0; it < iters; ++it) {
                    const T mid = T(0.5) * (lo + hi);
                    if (bez(g.type, g.y, mid) < ys) lo = mid; else hi = mid;
                }
                const T t = T(0.5) * (lo + hi);
                out.push_back({bez(g.type, g.x, t), t, (int)c, (int)j, dir, cpath[c]});
            }
        }
    }
}

inline bool filled(int rule, int w) { return rule ? (w & 1) != 0 : w != 0; }

template <class T> struct Model {
    int W, H, S, P;
    std::vector<Seg<T>> segs;
    std::vector<int> cpath, cvpos, rule;
    std::vector<Cross<T>> xs;           // sorted by (path, row, x)
    std::vector<int> dfill;             // per sorted crossing: fill left - fill right
    std::vector<T> cov;                 // [P][H][W]
    std::vector<T> pre;                 // [P][H][W][4] colour before each path
    std::vector<T> img;                 // [H][W][4]
};

template <class T>
void forward(Model<T> &m, const T *colors) {
    m.xs.clear();
    crossings(m.segs, m.cpath, m.H, m.S, m.xs);
    std::stable_sort(m.xs.begin(), m.xs.end(), [](const Cross<T> &a, const Cross<T> &b) {
        if (a.path != b.path) return a.path < b.path;
        if (a.row != b.row) return a.row < b.row;
        if (a.x != b.x) return a.x < b.x;
        if (a.curve != b.curve) return a.curve < b.curve;
        return a.t < b.t;
    });
    const size_t NX = m.xs.size();
    m.dfill.assign(NX, 0);
    m.cov.assign((size_t)m.P * m.H * m.W, T(0));
    for (size_t a = 0; a < NX;) {
        size_t b = a;
        while (b < NX && m.xs[b].path == m.xs[a].path && m.xs[b].row == m.xs[a].row) ++b;
        const int p = m.xs[a].path, j = m.xs[a].row, y = j / m.S, rule = m.rule[p];
        int w = 0;
        T prev = T(0);
        T *row = &m.cov[((size_t)p * m.H + y) * m.W];
        for (size_t k = a; k < b; ++k) {
            const T x = m.xs[k].x;
            if (filled(rule, w) && k > a) {
                // add [prev, x) clipped to the canvas
                const T u0 = std::max(prev, T(0)), u1 = std::min(x, T(m.W));
                if (u1 > u0) {
                    const long p0 = (long)std::floor(u0), p1 = std::min((long)std::floor(u1), (long)m.W - 1);
                    for (long px = p0; px <= p1; ++px) {
                        const T lo = std::max(u0, T(px)), hi = std::min(u1, T(px + 1));
                        if (hi > lo) row[px] += (hi - lo) / T(m.S);
                    }
                }
            }
            const int w2 = w + m.xs[k].dir;
            m.dfill[k] = (filled(rule, w) ? 1 : 0) - (filled(rule, w2) ? 1 : 0);
            w = w2;
            prev = x;
        }
        a = b;
    }
    const size_t NP = (size_t)m.W * m.H;
    m.pre.assign((size_t)m.P * NP * 4, T(0));
    m.img.assign(NP * 4, T(0));
    for (size_t i = 0; i < NP; ++i) {
        T I[4] = {T(1), T(1), T(1), T(1)};
        for (int p = 0; p < m.P; ++p) {
            T *pr = &m.pre[((size_t)p * NP + i) * 4];
            for (int ch = 0; ch < 4; ++ch) pr[ch] = I[ch];
            const T a = colors[4 * p + 3] * m.cov[(size_t)p * NP + i];
            const T oma = T(1) - a;
            for (int ch = 0; ch < 3; ++ch) I[ch] = colors[4 * p + ch] * a + I[ch] * oma;
            I[3] = a * a + I[3] * oma;
        }
        for (int ch = 0; ch < 4; ++ch) m.img[4 * i + ch] = I[ch];
    }
}

template <class T>
void backward(const Model<T> &m, const T *colors, const T *gimg, T *gpts, T *gcol) {
    const size_t NP = (size_t)m.W * m.H;
    std::vector<T> gcov((size_t)m.P * NP, T(0));
    for (int k = 0; k < 4 * m.P; ++k) gcol[k] = T(0);
    for (size_t i = 0; i < NP; ++i) {
        T g[4];
        for (int ch = 0; ch < 4; ++ch) g[ch] = gimg[4 * i + ch];
        for (int p = m.P - 1; p >= 0; --p) {
            const T *pr = &m.pre[((size_t)p * NP + i) * 4];
            const T cv = m.cov[(size_t)p * NP + i];
            const T pa = colors[4 * p + 3];
            const T a = pa * cv;
            T da = T(0);
            for (int ch = 0; ch < 3; ++ch) {
                da += g[ch] * (colors[4 * p + ch] - pr[ch]);
                gcol[4 * p + ch] += g[ch] * a;
            }
            da += g[3] * (T(2) * a - pr[3]);
            gcov[(size_t)p * NP + i] = da * pa;
            gcol[4 * p + 3] += da * cv;
            const T oma = T(1) - a;
            for (int ch = 0; ch < 4; ++ch) g[ch] *= oma;
        }
    }
    for (size_t k = 0; k < m.xs.size(); ++k) {
        const Cross<T> &c = m.xs[k];
        if (!m.dfill[k]) continue;
        if (!(c.x >= T(0) && c.x < T(m.W))) continue;
        const int y = c.row / m.S;
        const long px = (long)std::floor(c.x);
        const T gx = gcov[(size_t)c.path * NP + (size_t)y * m.W + px] * T(m.dfill[k]) / T(m.S);
        const Seg<T> &g = m.segs[c.curve];
        T b[4];
        basis(g.type, c.t, b);
        // dx/dy along the curve; near a y-extremum By' -> 0, bounded at |By'| >= 1e-4
        T dy = dbez(g.type, g.y, c.t);
        const T dx = dbez(g.type, g.x, c.t);
        if (std::fabs(dy) < T(1e-4)) dy = dy < T(0) ? T(-1e-4) : T(1e-4);
        const T ratio = dx / dy;
        const int nv = g.type;
        for (int v = 0; v < nv && v < 4; ++v) {
            const int vi = m.cvpos[c.curve] + v;
            gpts[2 * vi] += gx * b[v];
            gpts[2 * vi + 1] += -gx * ratio * b[v];
        }
    }
}

template <class T>
Model<T> *build(int W, int H, int S, int P, int C, const float *pts, const uint32_t *cvpos, const uint8_t *ctype,
                const uint32_t *cpath, const uint8_t *rule) {
    Model<T> *m = new Model<T>;
    m->W = W; m->H = H; m->S = S; m->P = P;
    m->segs.resize(C);
    m->cpath.resize(C);
    m->cvpos.resize(C);
    m->rule.assign(rule, rule + P);
    for (int c = 0; c < C; ++c) {
        Seg<T> &g = m->segs[c];
        g.type = ctype[c];
        const int nv = ctype[c] & 7;
        for (int k = 0; k < 4; ++k) {
            const int v = cvpos[c] + std::min(k, nv - 1);
            g.x[k] = (T)pts[2 * v];
            g.y[k] = (T)pts[2 * v + 1];
        }
        m->cpath[c] = (int)cpath[c];
        m->cvpos[c] = (int)cvpos[c];
    }
    return m;
}

// Outputs are written in the model's own precision (T).
template <class T>
int run(int W, int H, int S, int P, int C, int V, const float *pts, const uint32_t *cvpos, const uint8_t *ctype,
        const uint32_t *cpath, const uint8_t *rule, const float *colors, const T *gimg, T *img_out,
        T *gpts_out, T *gcol_out, int64_t *n_cross) {
    for (int c = 0; c < C; ++c)
        if (ctype[c] != 2 && ctype[c] != 3 && ctype[c] != 4) return 1;
    Model<T> *m = build<T>(W, H, S, P, C, pts, cvpos, ctype, cpath, rule);
    std::vector<T> col(colors, colors + 4 * P);
    forward(*m, col.data());
    if (img_out)
        for (size_t i = 0; i < m->img.size(); ++i) img_out[i] = m->img[i];
    if (n_cross) *n_cross = (int64_t)m->xs.size();
    if (gimg) {
        std::vector<T> g(gimg, gimg + 4 * (size_t)W * H), gp(2 * (size_t)V, T(0)), gc(4 * (size_t)P, T(0));
        backward(*m, col.data(), g.data(), gp.data(), gc.data());
        if (gpts_out)
            for (size_t i = 0; i < gp.size(); ++i) gpts_out[i] = gp[i];
        if (gcol_out)
            for (size_t i = 0; i < gc.size(); ++i) gcol_out[i] = gc[i];
    }
    delete m;
    return 0;
}

}  // namespace

extern "C" {

// Forward image (and, with gimg, the gradients of sum(gimg * image)) of the differentiable
// model.  precision 0 = float arithmetic (the device's), 1 = double; gimg, img_out,
// gpts_out and gcol_out are float or double arrays accordingly.  Returns nonzero for
// unsupported curve types (rational).
int oracle_diff_run(int precision, int W, int H, int S, int P, int C, int V, const float *pts,
                    const uint32_t *cvpos, const uint8_t *ctype, const uint32_t *cpath, const uint8_t *rule,
                    const float *colors, const void *gimg, void *img_out, void *gpts_out, void *gcol_out,
                    int64_t *n_cross) {
    if (precision)
        return run<double>(W, H, S, P, C, V, pts, cvpos, ctype, cpath, rule, colors, (const double *)gimg,
                           (double *)img_out, (double *)gpts_out, (double *)gcol_out, n_cross);
    return run<float>(W, H, S, P, C, V, pts, cvpos, ctype, cpath, rule, colors, (const float *)gimg,
                      (float *)img_out, (float *)gpts_out, (float *)gcol_out, n_cross);
}

}  // extern "C"
