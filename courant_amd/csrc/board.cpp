// Board model -> simulation front-end (host code of libcourant.so).
//
//   board model, KiCad import     web/src/import/kicad.ts:9-275, load.ts:20-29
//   copper geometry               web/src/import/kicad.ts:277-387
//   rule-based mesher             web/src/mesh/mesher.ts:47-156
//   board -> simulation           web/src/import/boardsim.ts:70-331
//
// The mesh decides every downstream number, so the mesher and the planner keep
// the reference's rules and their order of floating-point operations (the build
// has no FP contraction); the rest is organised for C++: a parsed S-expression
// tree, one Board value type, a Plan struct that the builder consumes, and
// exceptions inside that become cem_last_error() at the C boundary.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "courant.h"
#include "courant_board.h"
#include "sim.hpp"

namespace {

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kPi = 3.141592653589793;

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Pt {
    double x = 0, y = 0;
    double &at(int a) { return a == 0 ? x : y; }
    double at(int a) const { return a == 0 ? x : y; }
};

struct Layer {
    std::string name;
    int type = CEM_LAYER_DIELECTRIC;
    double thickness = 0;
    double eps_r = kNaN, loss_tangent = kNaN, roughness = kNaN;   // NaN: not given
    bool mask = false;
    double z0 = 0, z1 = 0;
};
struct Track {
    Pt a, b;
    double width = 0;
    std::string layer;
    int net = 0;
};
struct Via {
    Pt at;
    double drill = 0, size = 0;
    std::string la, lb;
    int net = 0;
};
struct Poly {
    std::vector<Pt> pts;
    std::string layer;
    int net = 0;
};

double or_default(double v, double d) { return std::isnan(v) ? d : v; }   // `v ?? d` for optional numbers
double truthy_or(double v, double d) { return (std::isnan(v) || v == 0) ? d : v; }   // `v || d`

}  // namespace

struct cem_board {
    std::vector<Layer> layers;                         // bottom to top
    std::vector<Track> tracks;
    std::vector<Via> vias;
    std::vector<Poly> polys;
    Pt omin, omax;                                     // bounding box of the outline
    std::vector<std::pair<int, std::string>> nets;     // first-set order, last name wins
    double thickness = 0;

    // z bottom-up: dielectrics stack, copper is a sheet at the surface (kicad.ts:130-135)
    void restack() {
        double z = 0;
        for (auto &l : layers) {
            if (l.type == CEM_LAYER_DIELECTRIC) {
                l.z0 = z;
                z += l.thickness;
                l.z1 = z;
            } else {
                l.z0 = z;
                l.z1 = z;
            }
        }
        thickness = z;
    }
    void set_net(int n, const std::string &name) {
        for (auto &e : nets)
            if (e.first == n) {
                e.second = name;
                return;
            }
        nets.push_back({n, name});
    }
    const Layer *layer(const std::string &name) const {
        for (auto &l : layers)
            if (l.name == name) return &l;
        return nullptr;
    }
    std::vector<const Layer *> of_type(int type) const {
        std::vector<const Layer *> out;
        for (auto &l : layers)
            if (l.type == type) out.push_back(&l);
        return out;
    }
};

namespace {

using Board = cem_board;

// ---------------------------------------------------------------------------
// copper geometry (kicad.ts:277-387)
// ---------------------------------------------------------------------------

// even-odd point in polygon, half-open crossing rule (kicad.ts:277-285)
bool in_poly(const Pt &p, const std::vector<Pt> &pts) {
    bool inside = false;
    const size_t n = pts.size();
    for (size_t i = 0, j = n ? n - 1 : 0; i < n; j = i++) {
        const double xi = pts[i].x, yi = pts[i].y, xj = pts[j].x, yj = pts[j].y;
        if ((yi > p.y) != (yj > p.y) && p.x < ((xj - xi) * (p.y - yi)) / (yj - yi) + xi) inside = !inside;
    }
    return inside;
}

// stadium of half-width w/2 around the segment (kicad.ts:287-295)
bool on_track(const Pt &p, const Track &t) {
    const double dx = t.b.x - t.a.x, dy = t.b.y - t.a.y, L2 = dx * dx + dy * dy;
    double u = L2 > 0 ? ((p.x - t.a.x) * dx + (p.y - t.a.y) * dy) / L2 : 0;
    u = std::max(0.0, std::min(1.0, u));
    const double qx = t.a.x + u * dx - p.x, qy = t.a.y + u * dy - p.y;
    const double hw = t.width / 2;
    return qx * qx + qy * qy <= hw * hw * (1 + 1e-9);
}

bool copper_at(const Board &b, const std::string &layer, const Pt &p) {
    for (auto &t : b.tracks)
        if (t.layer == layer && on_track(p, t)) return true;
    for (auto &q : b.polys)
        if (q.layer == layer && in_poly(p, q.pts)) return true;
    return false;
}

// Copper of one row y = const as x intervals that give copper_at's answers on the
// row: polygon crossings pair up half-open [lo, hi), track stadia are closed
// chords (end-cap circles and the body band clipped in x) (kicad.ts:312-361).
struct Row {
    std::vector<double> lo, hi;
    std::vector<char> closed;
    bool has(double x) const {
        for (size_t k = 0; k < lo.size(); k++)
            if (x >= lo[k] && (x < hi[k] || (closed[k] && x <= hi[k]))) return true;
        return false;
    }
};

Row copper_row(const Board &b, const std::string &layer, double y) {
    Row r;
    for (auto &t : b.tracks) {
        if (t.layer != layer) continue;
        const double hw0 = t.width / 2, h2 = hw0 * hw0 * (1 + 1e-9);
        double l = kInf, rr = -kInf;
        for (const Pt &c : {t.a, t.b}) {
            const double d = h2 - (y - c.y) * (y - c.y);
            if (d >= 0) {
                const double s = std::sqrt(d);
                l = std::min(l, c.x - s);
                rr = std::max(rr, c.x + s);
            }
        }
        const double ax = t.a.x, ay = t.a.y, dx = t.b.x - ax, dy = t.b.y - ay, L2 = dx * dx + dy * dy;
        if (L2 > 0) {
            const double L = std::sqrt(L2), hw = std::sqrt(h2);
            double lo = -kInf, hi = kInf;
            // keep mn <= k x + m <= mx
            auto clip = [&](double k, double m, double mn, double mx) {
                if (std::fabs(k) < 1e-300) {
                    if (m < mn || m > mx) {
                        lo = kInf;
                        hi = -kInf;
                    }
                    return;
                }
                const double x1 = (mn - m) / k, x2 = (mx - m) / k;
                lo = std::max(lo, std::min(x1, x2));
                hi = std::min(hi, std::max(x1, x2));
            };
            clip(dx / L2, (-ax * dx + (y - ay) * dy) / L2, 0, 1);      // along the segment
            clip(-dy / L, (ax * dy + (y - ay) * dx) / L, -hw, hw);    // across it
            if (lo <= hi) {
                l = std::min(l, lo);
                rr = std::max(rr, hi);
            }
        }
        if (l <= rr) {
            r.lo.push_back(l);
            r.hi.push_back(rr);
            r.closed.push_back(1);
        }
    }
    for (auto &q : b.polys) {
        if (q.layer != layer) continue;
        std::vector<double> xs;
        const size_t n = q.pts.size();
        for (size_t i = 0, j = n ? n - 1 : 0; i < n; j = i++) {
            const double xi = q.pts[i].x, yi = q.pts[i].y, xj = q.pts[j].x, yj = q.pts[j].y;
            if ((yi > y) != (yj > y)) xs.push_back(((xj - xi) * (y - yi)) / (yj - yi) + xi);
        }
        std::sort(xs.begin(), xs.end());
        for (size_t k = 0; k + 1 < xs.size(); k += 2) {
            r.lo.push_back(xs[k]);
            r.hi.push_back(xs[k + 1]);
            r.closed.push_back(0);
        }
    }
    return r;
}

// dedupe key of an edge: the coordinate to 1e-9 and the metal side (kicad.ts:365)
std::string edge_key(double x, int metal_hi) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.9f%c", x + 0.0, metal_hi ? 'h' : 'l');
    return buf;
}

struct Edge {
    double x;
    int metal_hi;   // 1: metal above x
};

// keeps the first position of each key with the last value written to it
struct EdgeSet {
    std::vector<Edge> items;
    std::vector<std::string> keys;
    void add(double x, int metal_hi) {
        const std::string k = edge_key(x, metal_hi);
        for (size_t i = 0; i < keys.size(); i++)
            if (keys[i] == k) {
                items[i] = Edge{x, metal_hi};
                return;
            }
        keys.push_back(k);
        items.push_back(Edge{x, metal_hi});
    }
};

// axis-aligned copper boundaries of constant coordinate `axis` with the metal side
// (kicad.ts:363-387): track sides of tracks running across `axis`, and polygon
// edges of constant `axis` (side from a point 1e-7 above the edge)
std::vector<Edge> copper_edges(const Board &b, const std::string &layer, int axis) {
    EdgeSet out;
    for (auto &t : b.tracks) {
        if (t.layer != layer) continue;
        const int along = std::fabs(t.b.x - t.a.x) > std::fabs(t.b.y - t.a.y) ? 0 : 1;
        if (along == axis) continue;
        const double c = t.a.at(axis);
        out.add(c - t.width / 2, 1);
        out.add(c + t.width / 2, 0);
    }
    for (auto &q : b.polys) {
        if (q.layer != layer) continue;
        const size_t n = q.pts.size();
        for (size_t i = 0; i < n; i++) {
            const Pt &p0 = q.pts[i], &p1 = q.pts[(i + 1) % n];
            const int o = 1 - axis;
            if (std::fabs(p0.at(axis) - p1.at(axis)) > 1e-12 || std::fabs(p0.at(o) - p1.at(o)) < 1e-12) continue;
            const double x = p0.at(axis);
            Pt mid;
            mid.at(o) = (p0.at(o) + p1.at(o)) / 2;
            mid.at(axis) = x + 1e-7;
            out.add(x, in_poly(mid, q.pts) ? 1 : 0);
        }
    }
    std::vector<Edge> v = out.items;
    std::stable_sort(v.begin(), v.end(), [](const Edge &a, const Edge &c) { return a.x < c.x; });
    return v;
}

// ---------------------------------------------------------------------------
// rule-based mesher (mesher.ts:47-156)
// ---------------------------------------------------------------------------
double lambda_over_20(double fmax, double er) { return 299792458.0 / (fmax * std::sqrt(er)) / 20; }

struct Span {
    double a, b;
    int n;
};
struct Cap {
    double a, b, h;
};
struct Features {
    double lo = 0, hi = 0;
    std::vector<Edge> sheets;
    std::vector<double> planes;
    std::vector<Span> spans;
    std::vector<Cap> caps;
    bool has_refine = false;            // false: refine at every plane and sheet edge
    std::vector<double> refine;
    std::vector<double> extra;
};
struct AxisMesh {
    std::vector<double> nodes;
    double min_cell = kInf, max_cell = 0, max_ratio = 1;
    std::vector<std::string> warnings;
};

std::string num_text(double v) {   // a coordinate in warning messages
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

// Fixed lines (domain ends, planes, thirds-rule pairs at sheet edges, extras),
// merged within hmin/4 by priority; then a size field of caps that relax with
// slope 0.8 (ratio - 1) away from where they apply, and each interval between
// fixed lines filled with ceil(integral dx / s) cells at equal steps of that
// integral (trapezoid rule on max(64, 8 (b - a) / hmin) sub-steps).
AxisMesh mesh_axis(const Features &f, double h0, double hmax, double ratio) {
    const double r = ratio > 0 ? ratio : 1.3;
    const double EPS = 1e-12;
    AxisMesh out;
    struct Line {
        double x;
        int pri;
    };
    std::vector<Line> fixed = {{f.lo, 3}, {f.hi, 3}};
    std::vector<double> refine = f.refine;
    for (double p : f.planes) {
        fixed.push_back({p, 2});
        if (!f.has_refine) refine.push_back(p);
    }
    for (const Edge &e : f.sheets) {
        const double s = e.metal_hi ? 1 : -1;   // direction into the metal
        fixed.push_back({e.x + s * h0 / 3, 1});
        fixed.push_back({e.x - s * 2 * h0 / 3, 1});
        if (!f.has_refine) refine.push_back(e.x);
    }
    for (double x : f.extra) fixed.push_back({x, 2});
    std::stable_sort(fixed.begin(), fixed.end(), [](const Line &a, const Line &b) { return a.x < b.x; });
    std::vector<double> lines;
    std::vector<int> pri;
    for (const Line &l : fixed) {
        if (l.x < f.lo - EPS || l.x > f.hi + EPS) continue;
        const size_t n = lines.size();
        if (n && l.x - lines[n - 1] < h0 / 4) {
            if (l.pri > pri[n - 1]) {
                lines[n - 1] = l.x;
                pri[n - 1] = l.pri;
            } else if (l.pri == pri[n - 1] && l.pri >= 2 && l.x - lines[n - 1] > EPS) {
                out.warnings.push_back("lines " + num_text(lines[n - 1]) + " and " + num_text(l.x) +
                                       " closer than hmin/4; merged");
            }
            continue;
        }
        lines.push_back(l.x);
        pri.push_back(l.pri);
    }
    const double g = 0.8 * (r - 1);
    auto dist = [](double x, double a, double b) { return x < a ? a - x : x > b ? x - b : 0.0; };
    auto size = [&](double x) {
        double s = hmax;
        for (const Cap &m : f.caps) s = std::min(s, m.h + g * dist(x, m.a, m.b));
        for (const Span &sp : f.spans) s = std::min(s, (sp.b - sp.a) / sp.n + g * dist(x, sp.a, sp.b));
        for (double q : refine) s = std::min(s, h0 + g * std::fabs(x - q));
        return std::max(s, h0 * 0.5);
    };
    if (lines.empty()) throw Fail("mesh axis: empty domain");
    out.nodes.push_back(lines[0]);
    std::vector<double> cum;
    for (size_t q = 0; q + 1 < lines.size(); q++) {
        const double a = lines[q], b = lines[q + 1];
        const double Md = std::ceil((b - a) / (h0 / 8));
        if (!(Md < 1e8)) throw Fail("mesh axis: hmin too small for the domain");
        const long M = std::max(64L, (long)Md);
        cum.assign(1, 0.0);
        for (long m = 0; m < M; m++) {
            const double x0 = a + ((b - a) * m) / M, x1 = a + ((b - a) * (m + 1)) / M;
            cum.push_back(cum[m] + (x1 - x0) * 0.5 * (1 / size(x0) + 1 / size(x1)));
        }
        const double I = cum[M];
        const long n = std::max(1L, (long)std::ceil(I - 1e-6));
        long m = 0;
        for (long k = 1; k < n; k++) {
            const double target = (I * k) / n;
            while (cum[m + 1] < target) m++;
            const double t = (target - cum[m]) / (cum[m + 1] - cum[m]);
            out.nodes.push_back(a + ((b - a) * (m + t)) / M);
        }
        out.nodes.push_back(b);
    }
    for (const Span &sp : f.spans) {
        long inside = 1;
        for (double v : out.nodes)
            if (v > sp.a + EPS && v < sp.b - EPS) inside++;
        if (inside < sp.n)
            out.warnings.push_back("span [" + num_text(sp.a) + ", " + num_text(sp.b) + "] has " + std::to_string(inside) +
                                   " cells, wanted " + std::to_string(sp.n));
    }
    for (size_t i = 0; i + 1 < out.nodes.size(); i++) {
        const double c = out.nodes[i + 1] - out.nodes[i];
        out.min_cell = std::min(out.min_cell, c);
        out.max_cell = std::max(out.max_cell, c);
        if (i > 0) {
            const double p = out.nodes[i] - out.nodes[i - 1];
            out.max_ratio = std::max(out.max_ratio, std::max(c / p, p / c));
        }
    }
    return out;
}

std::vector<double> with_pml(const std::vector<double> &nodes, int n, double h, bool lo, bool hi) {
    std::vector<double> out;
    if (lo)
        for (int i = n; i >= 1; i--) out.push_back(nodes[0] - i * h);
    out.insert(out.end(), nodes.begin(), nodes.end());
    if (hi)
        for (int i = 1; i <= n; i++) out.push_back(nodes.back() + i * h);
    return out;
}

// ---------------------------------------------------------------------------
// KiCad import (kicad.ts:12-275)
// ---------------------------------------------------------------------------
struct SExpr {
    bool list = false;
    std::string atom;
    std::vector<SExpr> items;
};

// lists, bare atoms and "quoted strings" (backslash takes the next character)
struct SReader {
    const char *s;
    size_t i = 0, n;
    int depth = 0;
    explicit SReader(const char *text) : s(text), n(std::strlen(text)) {}
    void skip() {
        while (i < n && std::isspace((unsigned char)s[i])) i++;
    }
    SExpr read() {
        skip();
        SExpr e;
        if (i < n && s[i] == '(') {
            if (++depth > 256) throw Fail("S-expression nested too deeply");
            i++;
            e.list = true;
            for (;;) {
                skip();
                if (i >= n) throw Fail("unbalanced parentheses");
                if (s[i] == ')') {
                    i++;
                    depth--;
                    return e;
                }
                e.items.push_back(read());
            }
        }
        if (i < n && s[i] == '"') {
            i++;
            while (i < n && s[i] != '"') {
                if (s[i] == '\\' && i + 1 < n) {
                    e.atom += s[i + 1];
                    i += 2;
                } else {
                    e.atom += s[i++];
                }
            }
            i++;
            return e;
        }
        const size_t start = i;
        while (i < n && !std::isspace((unsigned char)s[i]) && s[i] != '(' && s[i] != ')') i++;
        e.atom.assign(s + start, i - start);
        return e;
    }
};

const std::string &head(const SExpr &e) {
    static const std::string none;
    return e.list && !e.items.empty() && !e.items[0].list ? e.items[0].atom : none;
}
std::vector<const SExpr *> kids(const SExpr *e, const char *name) {
    std::vector<const SExpr *> out;
    if (e && e->list)
        for (auto &c : e->items)
            if (head(c) == name) out.push_back(&c);
    return out;
}
const SExpr *kid(const SExpr *e, const char *name) {
    if (e && e->list)
        for (auto &c : e->items)
            if (head(c) == name) return &c;
    return nullptr;
}
// a number as JavaScript's Number() reads an atom: trimmed, "" is 0, junk is NaN
double js_number(const std::string &t) {
    size_t a = 0, b = t.size();
    while (a < b && std::isspace((unsigned char)t[a])) a++;
    while (b > a && std::isspace((unsigned char)t[b - 1])) b--;
    if (a == b) return 0;
    const std::string s = t.substr(a, b - a);
    if (s == "Infinity" || s == "+Infinity") return kInf;
    if (s == "-Infinity") return -kInf;
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end != s.c_str() + s.size() || s.find_first_of("nNiI") != std::string::npos) return kNaN;
    return v;
}
double item_num(const SExpr *e, size_t k) {
    if (!e || !e->list || k >= e->items.size() || e->items[k].list) return kNaN;
    return js_number(e->items[k].atom);
}
double num(const SExpr *e, size_t k = 1) { return item_num(e, k); }
std::string str(const SExpr *e, size_t k = 1) {
    if (!e || !e->list || k >= e->items.size() || e->items[k].list) return "";
    return e->items[k].atom;
}

double to_m(double v) { return v * 1e-3; }
int to_int(double v) { return std::isfinite(v) && std::fabs(v) < 2e9 ? (int)v : 0; }   // net numbers
Pt P(double x, double y) { return Pt{to_m(x), -to_m(y)}; }   // KiCad mm, y down -> m, y up

Pt rotate(const Pt &p, double deg) {
    const double r = (deg * kPi) / 180, c = std::cos(r), s = std::sin(r);
    return Pt{p.x * c - p.y * s, p.x * s + p.y * c};
}

// points along the circular arc through p0, pm, p1, at most 10 degrees apart (kicad.ts:221-241)
std::vector<Pt> arc_through(const Pt &p0, const Pt &pm, const Pt &p1) {
    const double ax = p0.x, ay = p0.y, bx = pm.x, by = pm.y, cx = p1.x, cy = p1.y;
    const double d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by));
    if (std::fabs(d) < 1e-30) return {p0, p1};
    const double ux = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay) + (cx * cx + cy * cy) * (ay - by)) / d;
    const double uy = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx) + (cx * cx + cy * cy) * (bx - ax)) / d;
    const double r = std::hypot(ax - ux, ay - uy);
    const double t0 = std::atan2(ay - uy, ax - ux), tm = std::atan2(by - uy, bx - ux), t1 = std::atan2(cy - uy, cx - ux);
    auto norm = [](double t) {
        while (t < 0) t += 2 * kPi;
        return t;
    };
    double sw = norm(t1 - t0);
    if (norm(tm - t0) > sw) sw -= 2 * kPi;
    const long n = std::max(1L, (long)std::ceil(std::fabs(sw) / (kPi / 18)));
    std::vector<Pt> out = {p0};
    for (long k = 1; k < n; k++) {
        const double t = t0 + (sw * k) / n;
        out.push_back(Pt{ux + r * std::cos(t), uy + r * std::sin(t)});
    }
    out.push_back(p1);
    return out;
}

std::vector<Pt> hull(std::vector<Pt> p) {   // monotone chain (kicad.ts:268-275)
    std::sort(p.begin(), p.end(), [](const Pt &a, const Pt &b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
    auto cross = [](const Pt &o, const Pt &a, const Pt &b) { return (a.x - o.x) * (b.y - o.y) - (a.y - o.y) * (b.x - o.x); };
    std::vector<Pt> lo, hi;
    for (const Pt &q : p) {
        while (lo.size() >= 2 && cross(lo[lo.size() - 2], lo.back(), q) <= 0) lo.pop_back();
        lo.push_back(q);
    }
    for (size_t i = p.size(); i-- > 0;) {
        const Pt &q = p[i];
        while (hi.size() >= 2 && cross(hi[hi.size() - 2], hi.back(), q) <= 0) hi.pop_back();
        hi.push_back(q);
    }
    if (!lo.empty()) lo.pop_back();
    if (!hi.empty()) hi.pop_back();
    lo.insert(lo.end(), hi.begin(), hi.end());
    return lo;
}

// pad outline in pad-local KiCad units (mm, y down), before rotation (kicad.ts:244-266)
std::vector<Pt> pad_shape(const SExpr *pad, const std::string &shape, double w, double h) {
    auto circ = [](double cx, double cy, double r, double a0, double a1, int n) {
        std::vector<Pt> v;
        for (int k = 0; k <= n; k++)
            v.push_back(Pt{cx + r * std::cos(a0 + ((a1 - a0) * k) / n), cy + r * std::sin(a0 + ((a1 - a0) * k) / n)});
        return v;
    };
    auto rounded = [&](double r) {
        if (r <= 0) return std::vector<Pt>{{-w / 2, -h / 2}, {w / 2, -h / 2}, {w / 2, h / 2}, {-w / 2, h / 2}};
        const double ax = w / 2 - r, ay = h / 2 - r, q = kPi / 2;
        std::vector<Pt> v;
        for (auto part : {circ(ax, ay, r, 0, q, 6), circ(-ax, ay, r, q, 2 * q, 6), circ(-ax, -ay, r, 2 * q, 3 * q, 6),
                          circ(ax, -ay, r, 3 * q, 4 * q, 6)})
            v.insert(v.end(), part.begin(), part.end());
        return v;
    };
    if (shape == "circle") {
        auto v = circ(0, 0, w / 2, 0, 2 * kPi, 24);
        v.resize(24);
        return v;
    }
    if (shape == "oval") return rounded(std::min(w, h) / 2);
    if (shape == "roundrect") return rounded(truthy_or(num(kid(pad, "roundrect_rratio")), 0.25) * std::min(w, h));
    if (shape == "custom") {
        std::vector<Pt> all = rounded(0);
        if (const SExpr *prims = kid(pad, "primitives"))
            for (const SExpr *g : kids(prims, "gr_poly"))
                for (const SExpr *q : kids(kid(g, "pts"), "xy")) all.push_back(Pt{item_num(q, 1), item_num(q, 2)});
        return hull(all);
    }
    return rounded(0);   // rect, trapezoid and chamfered rects as rectangles
}

bool contains_ci(std::string s, const char *what) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s.find(what) != std::string::npos;
}

Board import_kicad(const char *text) {
    SReader rd(text);
    const SExpr root = rd.read();
    if (head(root) != "kicad_pcb") throw Fail("not a kicad_pcb file");
    Board b;
    for (const SExpr *n : kids(&root, "net")) b.set_net(to_int(item_num(n, 1)), str(n, 2));
    // stackup (setup (stackup (layer ...))), listed top-down
    std::vector<Layer> top_down;
    const SExpr *setup = kid(&root, "setup");
    if (const SExpr *stack = setup ? kid(setup, "stackup") : nullptr) {
        for (const SExpr *l : kids(stack, "layer")) {
            const std::string name = str(l, 1), type = str(kid(l, "type"));
            const double t = to_m(num(kid(l, "thickness")));
            const bool have_cu = std::any_of(top_down.begin(), top_down.end(),
                                             [](const Layer &q) { return q.type == CEM_LAYER_COPPER; });
            Layer L;
            L.name = name;
            if (type == "copper") {
                L.type = CEM_LAYER_COPPER;
                L.thickness = truthy_or(t, 35e-6);
            } else if (contains_ci(type, "solder mask") && !have_cu) {   // top mask only
                L.mask = true;
                L.thickness = truthy_or(t, 20e-6);
                L.eps_r = truthy_or(num(kid(l, "epsilon_r")), 3.5);
                L.loss_tangent = truthy_or(num(kid(l, "loss_tangent")), 0.02);
            } else if (type.find("core") != std::string::npos || type.find("prepreg") != std::string::npos) {
                L.thickness = t;
                L.eps_r = truthy_or(num(kid(l, "epsilon_r")), 4.5);
                L.loss_tangent = truthy_or(num(kid(l, "loss_tangent")), 0);
            } else {
                continue;
            }
            top_down.push_back(L);
        }
    }
    if (top_down.empty()) {   // no stackup: two copper layers on a 1.6 mm FR-4 core
        Layer f{"F.Cu", CEM_LAYER_COPPER, 35e-6}, core{"core", CEM_LAYER_DIELECTRIC, 1.6e-3 - 70e-6},
            bcu{"B.Cu", CEM_LAYER_COPPER, 35e-6};
        core.eps_r = 4.5;
        core.loss_tangent = 0.02;
        top_down = {f, core, bcu};
    }
    b.layers.assign(top_down.rbegin(), top_down.rend());
    b.restack();

    auto point = [](const SExpr *e) { return P(num(e, 1), num(e, 2)); };
    for (const SExpr *s : kids(&root, "segment"))
        b.tracks.push_back(Track{point(kid(s, "start")), point(kid(s, "end")), to_m(num(kid(s, "width"))),
                                 str(kid(s, "layer")), to_int(num(kid(s, "net")))});
    for (const SExpr *a : kids(&root, "arc")) {   // arc tracks (KiCad 7+) as short segments
        const auto pts = arc_through(point(kid(a, "start")), point(kid(a, "mid")), point(kid(a, "end")));
        for (size_t k = 0; k + 1 < pts.size(); k++)
            b.tracks.push_back(Track{pts[k], pts[k + 1], to_m(num(kid(a, "width"))), str(kid(a, "layer")),
                                     to_int(num(kid(a, "net")))});
    }
    for (const SExpr *v : kids(&root, "via")) {
        const SExpr *ly = kid(v, "layers");
        Via via;
        via.at = point(kid(v, "at"));
        via.size = to_m(num(kid(v, "size")));
        via.drill = to_m(num(kid(v, "drill")));
        via.la = ly && ly->items.size() > 1 ? ly->items[1].atom : "F.Cu";
        via.lb = ly && ly->items.size() > 2 ? ly->items[2].atom : "B.Cu";
        via.net = to_int(num(kid(v, "net")));
        b.vias.push_back(via);
    }
    for (const SExpr *zn : kids(&root, "zone")) {
        const int net = to_int(num(kid(zn, "net")));
        for (const SExpr *fp : kids(zn, "filled_polygon")) {
            Poly q;
            for (const SExpr *xy : kids(kid(fp, "pts"), "xy")) q.pts.push_back(P(item_num(xy, 1), item_num(xy, 2)));
            q.layer = str(kid(fp, "layer"));
            q.net = net;
            b.polys.push_back(q);
        }
    }
    const auto copper = b.of_type(CEM_LAYER_COPPER);
    for (const SExpr *fpn : kids(&root, "footprint")) {
        const SExpr *at = kid(fpn, "at");
        const double fx = num(at, 1), fy = num(at, 2), frot = truthy_or(num(at, 3), 0);
        for (const SExpr *pad : kids(fpn, "pad")) {
            const SExpr *pat = kid(pad, "at"), *sz = kid(pad, "size");
            const std::string shape = pad->items.size() > 3 && !pad->items[3].list ? pad->items[3].atom : "rect";
            const double prot = truthy_or(num(pat, 3), 0);
            const double w = num(sz, 1), h = num(sz, 2);
            const int net = to_int(num(kid(pad, "net")));
            // pad centre: the footprint rotation applies to the pad offset (KiCad, y down)
            const Pt off = rotate(Pt{num(pat, 1), -num(pat, 2)}, frot);
            const double cx = fx + off.x, cy = fy - off.y;
            std::vector<Pt> pts;
            for (const Pt &q : pad_shape(pad, shape, w, h)) {
                const Pt r = rotate(Pt{q.x, -q.y}, prot);
                pts.push_back(P(cx + r.x, cy - r.y));
            }
            if (const SExpr *ly = kid(pad, "layers")) {
                for (size_t k = 1; k < ly->items.size(); k++) {
                    const std::string &lname = ly->items[k].atom;
                    if (lname == "*.Cu") {
                        for (const Layer *c : copper) b.polys.push_back(Poly{pts, c->name, net});
                    } else if (lname.size() >= 3 && lname.compare(lname.size() - 3, 3, ".Cu") == 0) {
                        b.polys.push_back(Poly{pts, lname, net});
                    }
                }
            }
            // a plated through-hole pad gets a barrel, like a via
            const SExpr *drill = kid(pad, "drill");
            if (pad->items.size() > 2 && pad->items[2].atom == "thru_hole" && drill && !copper.empty()) {
                double d;
                if (str(drill, 1) == "oval") {
                    const double d2 = item_num(drill, 2);
                    const double d3 = drill->items.size() > 3 ? item_num(drill, 3) : d2;
                    d = std::min(d2, d3);
                } else {
                    d = item_num(drill, 1);
                }
                b.vias.push_back(Via{P(cx, cy), to_m(d), to_m(std::min(w, h)), copper.back()->name, copper.front()->name, net});
            }
        }
    }
    // outline: bounding box of the Edge.Cuts graphics
    double x0 = kInf, y0 = kInf, x1 = -kInf, y1 = -kInf;
    auto grow = [&](const Pt &p) {
        x0 = std::min(x0, p.x);
        x1 = std::max(x1, p.x);
        y0 = std::min(y0, p.y);
        y1 = std::max(y1, p.y);
    };
    for (const char *kind : {"gr_line", "gr_rect"})
        for (const SExpr *g : kids(&root, kind)) {
            if (str(kid(g, "layer")) != "Edge.Cuts") continue;
            grow(point(kid(g, "start")));
            grow(point(kid(g, "end")));
        }
    for (const SExpr *g : kids(&root, "gr_arc")) {
        if (str(kid(g, "layer")) != "Edge.Cuts" || !kid(g, "mid")) continue;
        for (const Pt &p : arc_through(point(kid(g, "start")), point(kid(g, "mid")), point(kid(g, "end")))) grow(p);
    }
    for (const SExpr *g : kids(&root, "gr_circle")) {
        if (str(kid(g, "layer")) != "Edge.Cuts") continue;
        const Pt c = point(kid(g, "center")), e = point(kid(g, "end"));
        const double r = std::hypot(e.x - c.x, e.y - c.y);
        grow(Pt{c.x - r, c.y - r});
        grow(Pt{c.x + r, c.y + r});
    }
    for (const SExpr *g : kids(&root, "gr_poly")) {
        if (str(kid(g, "layer")) != "Edge.Cuts") continue;
        for (const SExpr *q : kids(kid(g, "pts"), "xy")) grow(P(item_num(q, 1), item_num(q, 2)));
    }
    if (!std::isfinite(x0)) throw Fail("board outline (Edge.Cuts) not found");
    b.omin = Pt{x0, y0};
    b.omax = Pt{x1, y1};
    return b;
}

// ---------------------------------------------------------------------------
// board -> simulation (boardsim.ts:70-331)
// ---------------------------------------------------------------------------
int nearest(const std::vector<double> &a, double v) {
    int b = 0;
    for (int i = 1; i < (int)a.size(); i++)
        if (std::fabs(a[i] - v) < std::fabs(a[b] - v)) b = i;
    return b;
}

struct PortSpec {
    bool by_name;
    std::string name;
    int net;
    int edge;
    double offset;
};

int net_number(const Board &b, const PortSpec &p) {
    if (!p.by_name) return p.net;
    for (auto &e : b.nets)
        if (e.second == p.name) return e.first;
    throw Fail("net " + p.name + " not found");
}

// an x-directed track of `net` that reaches or crosses the plane x = xc (boardsim.ts:90-94)
const Track *track_across(const Board &b, int net, double xc) {
    for (auto &t : b.tracks)
        if (t.net == net && std::fabs(t.a.y - t.b.y) < 1e-9 && std::min(t.a.x, t.b.x) <= xc + 1e-6 &&
            std::max(t.a.x, t.b.x) >= xc - 1e-6)
            return &t;
    return nullptr;
}

// quasi-static effective index of a strip of width w at height z0 (boardsim.ts:81-88)
double neff_guess(const Board &b, double z0, double w) {
    double h = 0, erh = 0;
    int n = 0;
    for (auto &l : b.layers) {
        if (l.type != CEM_LAYER_DIELECTRIC || l.mask || !(l.z1 <= z0 + 1e-12)) continue;
        h += l.thickness;
        erh += or_default(l.eps_r, 4.5) * l.thickness;
        n++;
    }
    h = std::max(1e-6, h);
    const double er = n ? erh / h : 1;
    const double ee = (er + 1) / 2 + ((er - 1) / 2) / std::sqrt(1 + 12 * h / w);
    return std::sqrt(ee);
}

struct Options {
    double fmax, df, hmin, feed, mon, air_above, air_side, y_window;
    int npml, excite;
    bool lossy, thick;
};

Options options_of(const cem_board_options *o) {
    if (!o) throw Fail("board options missing");
    if (!(o->fmax > 0) || !(o->df > 0) || !(o->hmin > 0)) throw Fail("board options need fmax, df and hmin > 0");
    Options r;
    r.fmax = o->fmax;
    r.df = o->df;
    r.hmin = o->hmin;
    r.feed = or_default(o->feed, 4e-3);
    r.mon = or_default(o->monitor, 2.5e-3);
    r.air_above = or_default(o->air_above, 1.5e-3);
    r.air_side = or_default(o->air_side, 1e-3);
    r.y_window = o->y_window > 0 ? o->y_window : 0;
    r.npml = o->npml < 0 ? 8 : o->npml;
    r.excite = o->excite;
    r.lossy = o->lossy != 0;
    r.thick = o->thick_copper != 0;
    return r;
}

std::vector<PortSpec> ports_of(const cem_board_port *ports, int n) {
    std::vector<PortSpec> v;
    for (int k = 0; k < n; k++) {
        const cem_board_port &p = ports[k];
        if (p.edge != CEM_PORT_XLO && p.edge != CEM_PORT_XHI) throw Fail("port edge must be xlo or xhi");
        v.push_back(PortSpec{p.net_name != nullptr, p.net_name ? p.net_name : "", p.net, p.edge, p.offset});
    }
    return v;
}

// the board restricted to a y band around the port tracks (boardsim.ts:63-79)
Board crop_to_net(const Board &b, const Options &o, const std::vector<PortSpec> &ports) {
    std::vector<double> ys;
    for (auto &ps : ports) {
        const double plane = ps.edge == CEM_PORT_XLO ? b.omin.x + ps.offset : b.omax.x - ps.offset;
        if (const Track *t = track_across(b, net_number(b, ps), plane)) {
            ys.push_back(t->a.y - t->width / 2);
            ys.push_back(t->a.y + t->width / 2);
        }
    }
    if (ys.empty()) return b;
    const double lo = std::max(b.omin.y, *std::min_element(ys.begin(), ys.end()) - o.y_window);
    const double hi = std::min(b.omax.y, *std::max_element(ys.begin(), ys.end()) + o.y_window);
    auto y_in = [&](double a, double c) { return c >= lo && a <= hi; };
    Board c = b;
    c.tracks.clear();
    c.polys.clear();
    c.vias.clear();
    for (auto &t : b.tracks)
        if (y_in(std::min(t.a.y, t.b.y) - t.width / 2, std::max(t.a.y, t.b.y) + t.width / 2)) c.tracks.push_back(t);
    for (auto &q : b.polys) {
        double a = kInf, e = -kInf;
        for (auto &p : q.pts) {
            a = std::min(a, p.y);
            e = std::max(e, p.y);
        }
        if (y_in(a, e)) c.polys.push_back(q);
    }
    for (auto &v : b.vias)
        if (v.at.y >= lo && v.at.y <= hi) c.vias.push_back(v);
    c.omin = Pt{b.omin.x, lo};
    c.omax = Pt{b.omax.x, hi};
    return c;
}

// mesh and geometry decisions of a board run (boardsim.ts:110-196)
struct Plan {
    Board b;
    Options o;
    std::vector<PortSpec> ports;
    double bx0, bx1, by0, by1, cx0, cx1, X0, X1, yLo, yHi, zTop, hp;
    bool has_lo, has_hi;
    std::vector<const Layer *> dielectrics, copper;
    std::vector<Via> inside;
    std::vector<double> x, y, z;
    int nx, ny, nz;
    double clamp_x(double v) const { return std::min(cx1, std::max(cx0, v)); }
    double clamp_y(double v) const { return std::min(by1, std::max(by0, v)); }
    double cu_top(const Layer &l) const { return l.z0 + (o.thick ? l.thickness : 0); }
};

std::unique_ptr<Plan> plan_board(const Board &board, const Options &o, const std::vector<PortSpec> &ports) {
    auto P_ = std::make_unique<Plan>();
    Plan &p = *P_;
    p.o = o;
    p.ports = ports;
    p.b = o.y_window > 0 ? crop_to_net(board, o, ports) : board;
    const Board &b = p.b;
    p.bx0 = b.omin.x;
    p.by0 = b.omin.y;
    p.bx1 = b.omax.x;
    p.by1 = b.omax.y;
    // where a band cut through the board, the cross-section runs into the y CPML
    const bool cut_lo = p.by0 > board.omin.y + 1e-9, cut_hi = p.by1 < board.omax.y - 1e-9;
    p.yLo = cut_lo ? -kInf : p.by0;
    p.yHi = cut_hi ? kInf : p.by1;
    p.dielectrics = b.of_type(CEM_LAYER_DIELECTRIC);
    p.copper = b.of_type(CEM_LAYER_COPPER);
    if (p.copper.empty()) throw Fail("board has no copper layers");
    double er_max = 1;
    for (const Layer *l : p.dielectrics) er_max = std::max(er_max, or_default(l->eps_r, 1));
    const double h_air = std::min(lambda_over_20(o.fmax, 1), 200e-6);
    const double h_diel = lambda_over_20(o.fmax, er_max);
    const Layer &bottom = *p.copper.front();
    {   // the bottom copper is the domain's PEC floor if it covers the board
        int hits = 0, tot = 0;
        for (int i = 0; i < 20; i++)
            for (int j = 0; j < 20; j++) {
                tot++;
                const Pt q{p.bx0 + (p.bx1 - p.bx0) * (i + 0.5) / 20, p.by0 + (p.by1 - p.by0) * (j + 0.5) / 20};
                if (copper_at(b, bottom.name, q)) hits++;
            }
        if (!((double)hits / tot > 0.95)) throw Fail("milestone 1 needs a solid bottom ground plane");
    }
    p.has_lo = p.has_hi = false;
    double off_lo = 0, off_hi = 0;
    bool seen_lo = false, seen_hi = false;
    for (auto &ps : ports) {
        if (ps.edge == CEM_PORT_XLO) {
            p.has_lo = true;
            if (!seen_lo) off_lo = ps.offset, seen_lo = true;
        } else {
            p.has_hi = true;
            if (!seen_hi) off_hi = ps.offset, seen_hi = true;
        }
    }
    // cut planes: the board between them is simulated, the cross-section at a plane is extruded
    p.cx0 = p.bx0 + (p.has_lo ? off_lo : 0);
    p.cx1 = p.bx1 - (p.has_hi ? off_hi : 0);
    if (!(p.cx1 - p.cx0 > 1e-4)) throw Fail("port planes leave less than 0.1 mm of board");
    p.X0 = p.cx0 - (p.has_lo ? o.feed : 0);
    p.X1 = p.cx1 + (p.has_hi ? o.feed : 0);

    // ---- mesh ----
    const Layer &top = *p.copper.back();
    p.zTop = top.z1 + (o.thick ? top.thickness : 0);
    std::vector<double> via_x, via_y;
    for (auto &v : b.vias)
        if (v.at.x > p.cx0 && v.at.x < p.cx1) {   // vias beyond the port planes are cut away
            p.inside.push_back(v);
            via_x.push_back(v.at.x);
            via_y.push_back(v.at.y);
        }
    EdgeSet sheet_y, sheet_x;   // edges of all copper layers, duplicates merged
    for (const Layer *c : p.copper) {
        for (const Edge &e : copper_edges(b, c->name, 1))
            if (e.x > p.by0 - 1e-9 && e.x < p.by1 + 1e-9) sheet_y.add(e.x, e.metal_hi);
        for (const Edge &e : copper_edges(b, c->name, 0))
            if (e.x > p.cx0 + 1e-9 && e.x < p.cx1 - 1e-9) sheet_x.add(e.x, e.metal_hi);
    }
    double via_r = 0;
    if (!p.inside.empty()) {
        via_r = kInf;
        for (auto &v : p.inside) via_r = std::min(via_r, v.drill / 2);
    }
    const double h_via = via_r ? std::max(o.hmin, via_r / 2) : kInf;

    Features fy;
    fy.lo = cut_lo ? p.by0 : p.by0 - o.air_side;
    fy.hi = cut_hi ? p.by1 : p.by1 + o.air_side;
    fy.sheets = sheet_y.items;
    fy.planes = {p.by0, p.by1};
    fy.has_refine = true;
    for (auto &e : sheet_y.items) fy.refine.push_back(e.x);
    fy.refine.insert(fy.refine.end(), via_y.begin(), via_y.end());
    fy.caps.push_back(Cap{p.by0, p.by1, h_diel});
    for (double v : via_y) fy.caps.push_back(Cap{v - via_r, v + via_r, h_via});
    const AxisMesh ym = mesh_axis(fy, o.hmin, h_air, 1.3);

    Features fx;
    fx.lo = p.X0;
    fx.hi = p.X1;
    fx.sheets = sheet_x.items;
    fx.planes = {p.cx0, p.cx1};
    fx.has_refine = true;
    for (auto &e : sheet_x.items) fx.refine.push_back(e.x);
    fx.refine.insert(fx.refine.end(), via_x.begin(), via_x.end());
    fx.caps.push_back(Cap{p.X0, p.X1, std::min(h_diel, 4 * o.hmin)});
    for (double v : via_x) fx.caps.push_back(Cap{v - via_r, v + via_r, h_via});
    const AxisMesh xm = mesh_axis(fx, o.hmin, h_air, 1.3);

    Features fz;
    fz.lo = 0;
    fz.hi = p.zTop + o.air_above;
    {
        std::vector<double> zp;
        auto add = [&](double v) {
            if (std::find(zp.begin(), zp.end(), v) == zp.end()) zp.push_back(v);
        };
        for (auto &l : b.layers) {
            add(l.z0);
            add(l.z1);
        }
        for (const Layer *c : p.copper)
            if (c != &bottom) add(p.cu_top(*c));
        fz.planes = zp;
    }
    for (const Layer *l : p.dielectrics) {
        fz.spans.push_back(Span{l->z0, l->z1, l->mask ? 1 : 4});
        fz.caps.push_back(Cap{l->z0, l->z1, h_diel});
    }
    fz.has_refine = true;
    for (const Layer *c : p.copper) fz.refine.push_back(c->z0);
    if (o.thick)
        for (const Layer *c : p.copper)
            if (c != &bottom) fz.refine.push_back(p.cu_top(*c));
    const AxisMesh zm = mesh_axis(fz, o.hmin, h_air, 1.3);

    p.hp = h_air / 2;
    p.x = with_pml(xm.nodes, o.npml, p.hp, p.has_lo, p.has_hi);
    p.y = with_pml(ym.nodes, o.npml, p.hp, true, true);
    p.z = with_pml(zm.nodes, o.npml, p.hp, false, true);
    p.nx = (int)p.x.size() - 1;
    p.ny = (int)p.y.size() - 1;
    p.nz = (int)p.z.size() - 1;
    return P_;
}

void check(int rc, const char *what) {
    if (rc < 0) throw Fail(std::string(what) + ": " + cem_last_error());
}

double layer_z0(const Board &b, const std::string &name, double fallback) {
    const Layer *l = b.layer(name);
    return l ? l->z0 : fallback;
}

// the finalized simulation of a plan (boardsim.ts:198-311)
void build_into(cem_sim *sim, const Plan &p, int *port_ids, double *t_stop) {
    const Board &b = p.b;
    const Options &o = p.o;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const std::vector<double> &x = p.x, &y = p.y, &z = p.z;
    if (p.has_lo) check(cem_set_boundary(sim, CEM_XLO, CEM_BC_CPML, o.npml), "set_boundary");
    if (p.has_hi) check(cem_set_boundary(sim, CEM_XHI, CEM_BC_CPML, o.npml), "set_boundary");
    check(cem_set_boundary(sim, CEM_YLO, CEM_BC_CPML, o.npml), "set_boundary");
    check(cem_set_boundary(sim, CEM_YHI, CEM_BC_CPML, o.npml), "set_boundary");
    check(cem_set_boundary(sim, CEM_ZHI, CEM_BC_CPML, o.npml), "set_boundary");
    const double big = 1;
    const double ylo = std::max(-big, p.yLo), yhi = std::min(big, p.yHi);

    // dielectrics, extruded in x into the CPML
    if (o.lossy) {
        if (p.dielectrics.empty()) throw Fail("lossy board needs a dielectric layer");
        const int np = 8;
        std::vector<std::vector<double>> deps;
        std::vector<double> einf;
        std::vector<double> fp0(np);
        for (size_t i = 0; i < p.dielectrics.size(); i++) {
            const Layer &l = *p.dielectrics[i];
            std::vector<double> fp(np), d(np);
            double e = 0;
            cem_fit_djordjevic_sarkar(or_default(l.eps_r, 4.5), or_default(l.loss_tangent, 0), 10e9, o.df, o.fmax, np, &e,
                                      fp.data(), d.data());
            if (i == 0) fp0 = fp;   // the pole frequencies depend only on the band
            einf.push_back(e);
            deps.push_back(d);
        }
        check(cem_set_debye_poles(sim, np, fp0.data()), "set_debye_poles");
        for (size_t i = 0; i < p.dielectrics.size(); i++) {
            const Layer &l = *p.dielectrics[i];
            check(cem_material_box_debye(sim, -big, big, ylo, yhi, l.z0, l.z1, einf[i], 0, deps[i].data()),
                  "material_box_debye");
        }
        double rq = 0;   // roughness of the port tracks' copper (one SIBC model for the board)
        for (const Layer *c : p.copper) rq = std::max(rq, or_default(c->roughness, 0));
        check(cem_set_conductor_loss(sim, 5.8e7, rq > 0 ? 1 : 0, rq, 0, o.df, o.fmax), "set_conductor_loss");
    } else {
        for (const Layer *l : p.dielectrics)
            check(cem_material_box(sim, -big, big, ylo, yhi, l->z0, l->z1, or_default(l->eps_r, 4.5), 0), "material_box");
    }

    // copper sheets rasterised edge by edge (the bottom layer is the PEC floor): an
    // edge is metal if both ends and its midpoint are on copper, sampled row by row
    const Layer *bottom = p.copper.front();
    std::vector<double> xc(nx + 1), xm(nx);
    for (int i = 0; i <= nx; i++) xc[i] = p.clamp_x(x[i]);
    for (int i = 0; i < nx; i++) xm[i] = p.clamp_x((x[i] + x[i + 1]) / 2);
    for (const Layer *c : p.copper) {
        if (c == bottom) continue;
        const int k = nearest(z, c->z0);
        std::vector<unsigned char> mu((size_t)nx * (ny + 1), 0), mv((size_t)(nx + 1) * ny, 0);
        std::vector<std::vector<unsigned char>> node(ny + 1, std::vector<unsigned char>(nx + 1, 0));
        for (int j = 0; j <= ny; j++) {
            if (y[j] < p.yLo - 1e-12 || y[j] > p.yHi + 1e-12) continue;
            const Row r = copper_row(b, c->name, p.clamp_y(y[j]));
            for (int i = 0; i <= nx; i++) node[j][i] = r.has(xc[i]);
            for (int i = 0; i < nx; i++)
                if (node[j][i] && node[j][i + 1] && r.has(xm[i])) mu[(size_t)i * (ny + 1) + j] = 1;
        }
        for (int j = 0; j < ny; j++) {
            if (y[j] < p.yLo - 1e-12 || y[j + 1] > p.yHi + 1e-12) continue;
            const Row r = copper_row(b, c->name, p.clamp_y((y[j] + y[j + 1]) / 2));
            for (int i = 0; i <= nx; i++)
                if (node[j][i] && node[j + 1][i] && r.has(xc[i])) mv[(size_t)i * ny + j] = 1;
        }
        const int kt = o.thick ? nearest(z, p.cu_top(*c)) : k;
        const int nk = kt - k;
        if (!o.thick || nk <= 0) {
            check(cem_pec_edges_plane(sim, CEM_Z, k, mu.data(), mv.data()), "pec_edges_plane");
            continue;
        }
        // thick copper: the cells between the layer surface and its top face whose
        // centre is on copper
        std::vector<unsigned char> cells((size_t)nx * ny * nk, 0);
        for (int j = 0; j < ny; j++) {
            const double yc = (y[j] + y[j + 1]) / 2;
            if (yc < p.yLo || yc > p.yHi) continue;
            const Row r = copper_row(b, c->name, p.clamp_y(yc));
            for (int i = 0; i < nx; i++)
                if (r.has(xm[i])) std::fill_n(cells.begin() + ((size_t)i * ny + j) * nk, nk, 1);
        }
        check(cem_pec_cells(sim, 0, 0, k, nx, ny, nk, cells.data()), "pec_cells");
    }

    // via barrels: solid PEC cells inside the drill radius, between their layers
    auto via_span = [&](const Via &v, double *zl, double *zh) {
        const double a = layer_z0(b, v.la, 0), c = layer_z0(b, v.lb, 0);
        const double a2 = layer_z0(b, v.la, p.zTop), c2 = layer_z0(b, v.lb, p.zTop);
        *zl = std::min(a, c);
        *zh = std::max(a2, c2);
    };
    for (const Via &v : p.inside) {
        double zl, zh;
        via_span(v, &zl, &zh);
        const double r = v.drill / 2;
        const int i0 = std::max(0, nearest(x, v.at.x - r) - 1), i1 = std::min(nx - 1, nearest(x, v.at.x + r) + 1);
        const int j0 = std::max(0, nearest(y, v.at.y - r) - 1), j1 = std::min(ny - 1, nearest(y, v.at.y + r) + 1);
        const int k0 = nearest(z, zl), k1 = nearest(z, zh);
        const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0;
        if (nk <= 0 || ni <= 0 || nj <= 0) continue;
        std::vector<unsigned char> m((size_t)ni * nj * nk, 0);
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++) {
                const double cx = (x[i0 + i] + x[i0 + i + 1]) / 2, cy = (y[j0 + j] + y[j0 + j + 1]) / 2;
                const double dx = cx - v.at.x, dy = cy - v.at.y;
                if (dx * dx + dy * dy <= r * r) std::fill_n(m.begin() + ((size_t)i * nj + j) * nk, nk, 1);
            }
        check(cem_pec_cells(sim, i0, j0, k0, ni, nj, nk, m.data()), "pec_cells");
    }

    // feed extensions: the ground vias near a port plane become continuous walls
    // through the feed, between their layers, so the feed's top grounds do not float
    int floor_net = 0;
    bool have_floor_net = false;
    {
        std::vector<std::pair<int, long>> count;   // first-seen order
        for (auto &q : b.polys) {
            if (q.layer != bottom->name) continue;
            auto it = std::find_if(count.begin(), count.end(), [&](const auto &e) { return e.first == q.net; });
            if (it == count.end()) count.push_back({q.net, (long)q.pts.size()});
            else it->second += (long)q.pts.size();
        }
        std::stable_sort(count.begin(), count.end(), [](const auto &a, const auto &c) { return a.second > c.second; });
        if (!count.empty()) floor_net = count[0].first, have_floor_net = true;
    }
    auto feed_walls = [&](double xa, double xb, double plane) {
        const int i0 = nearest(x, xa), i1 = nearest(x, xb);
        if (i1 <= i0) return;
        std::set<std::string> seen;
        for (const Via &v : p.inside) {
            if (std::fabs(v.at.x - plane) > 2e-3 || (have_floor_net && v.net != floor_net && v.net != 0)) continue;
            char key[64];
            std::snprintf(key, sizeof key, "%.7f", v.at.y + 0.0);
            if (!seen.insert(key).second) continue;
            double zl, zh;
            via_span(v, &zl, &zh);
            const double r = v.drill / 2;
            check(cem_pec_box(sim, x[i0], x[i1], v.at.y - r, v.at.y + r, zl, zh), "pec_box");
        }
    };
    if (p.has_lo) feed_walls(p.X0 - o.npml * p.hp, p.cx0, p.cx0);
    if (p.has_hi) feed_walls(p.cx1, p.X1 + o.npml * p.hp, p.cx1);

    // mode ports where the port tracks cross the port planes
    const double tau = std::sqrt(6.9) / (kPi * 1.25 * o.fmax);
    for (size_t q = 0; q < p.ports.size(); q++) {
        const PortSpec &ps = p.ports[q];
        const int net = net_number(b, ps);
        const double edge_x = ps.edge == CEM_PORT_XLO ? p.cx0 : p.cx1;
        const Track *t = track_across(b, net, edge_x);
        if (!t)
            throw Fail("no x-directed track of net " + (ps.by_name ? ps.name : std::to_string(ps.net)) + " crosses the " +
                       (ps.edge == CEM_PORT_XLO ? "xlo" : "xhi") + " port plane");
        const Layer *layer = b.layer(t->layer);
        if (!layer) throw Fail("port track on unknown layer " + t->layer);
        const double yc = t->a.y, hw = t->width / 2;
        const int kh = nearest(z, layer->z0);
        const int jl = nearest(y, yc - hw), jr = nearest(y, yc + hw);
        // current loop on dual nodes one cell outside each strip edge
        const int ja = std::max(0, (y[jl] <= yc - hw ? jl : jl - 1) - 1);
        const int jb = std::min(ny - 1, (y[jr] >= yc + hw ? jr : jr + 1));
        const bool lo = ps.edge == CEM_PORT_XLO;
        const int i_src = lo ? o.npml + 2 : nx - o.npml - 2;
        const int i_mon = lo ? nearest(x, x[i_src] + o.mon) : nearest(x, x[i_src] - o.mon) - 1;
        cem_mode_port_def d{};
        d.dir = lo ? 1 : -1;
        d.i_src = i_src;
        d.i_mon = i_mon;
        d.j0 = 0;
        d.j1 = ny;
        d.k0 = 0;
        d.k1 = nz;
        d.vj = nearest(y, yc);
        d.vk0 = 0;
        d.vk1 = kh;
        d.vsign = -1;
        d.ja = ja;
        d.jb = jb;
        d.ka = kh - 1;
        d.kb = o.thick ? std::max(kh, nearest(z, p.cu_top(*layer))) : kh;
        d.x_ref = edge_x;
        d.f0 = 0.5 * o.fmax;
        d.neff0 = neff_guess(b, layer->z0, t->width);
        port_ids[q] = cem_add_mode_port(sim, &d);
        check(port_ids[q], "add_mode_port");
    }
    if (o.excite < 0 || o.excite >= (int)p.ports.size()) throw Fail("excited port out of range");
    check(cem_set_mode_excitation(sim, port_ids[o.excite], 4.5 * tau, tau), "set_mode_excitation");
    const long nf = std::lround(std::floor(o.fmax / o.df + 0.5));
    std::vector<double> freqs(nf);
    for (long i = 0; i < nf; i++) freqs[i] = (i + 1) * o.df;
    check(cem_set_frequencies(sim, (int)nf, freqs.data()), "set_frequencies");
    if (cem_finalize(sim, 0.99) != 0) throw Fail(std::string("finalize: ") + cem_last_error());
    if (t_stop) *t_stop = 10.5 * tau;
}

void copy_name(const std::string &s, char *out, int cap) {
    if (!out || cap <= 0) return;
    std::snprintf(out, (size_t)cap, "%s", s.c_str());
}

template <class F>
int guarded(F fn) {
    try {
        return fn();
    } catch (const std::exception &e) {
        cem::set_error(e.what());
        return -1;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" cem_board *cem_board_create(void) { return new (std::nothrow) cem_board(); }

extern "C" cem_board *cem_board_import_kicad(const char *text) {
    if (!text) {
        cem::set_error("no text");
        return nullptr;
    }
    try {
        return new cem_board(import_kicad(text));
    } catch (const std::exception &e) {
        cem::set_error(e.what());
        return nullptr;
    }
}

extern "C" void cem_board_destroy(cem_board *b) { delete b; }

extern "C" cem_board *cem_board_clone(const cem_board *b) { return b ? new (std::nothrow) cem_board(*b) : nullptr; }

extern "C" int cem_board_add_layer(cem_board *b, const char *name, int type, double thickness, double eps_r,
                                   double loss_tangent, double roughness, int mask) {
    if (!b || !name || (type != CEM_LAYER_COPPER && type != CEM_LAYER_DIELECTRIC)) {
        cem::set_error("bad layer");
        return -1;
    }
    Layer l;
    l.name = name;
    l.type = type;
    l.thickness = thickness;
    l.eps_r = eps_r;
    l.loss_tangent = loss_tangent;
    l.roughness = roughness;
    l.mask = mask != 0;
    b->layers.push_back(l);
    b->restack();
    return (int)b->layers.size() - 1;
}

extern "C" int cem_board_clear_layers(cem_board *b) {
    if (!b) return -1;
    b->layers.clear();
    b->restack();
    return 0;
}

extern "C" int cem_board_add_track(cem_board *b, double ax, double ay, double bx, double by, double width,
                                   const char *layer, int net) {
    if (!b || !layer) return -1;
    b->tracks.push_back(Track{Pt{ax, ay}, Pt{bx, by}, width, layer, net});
    return (int)b->tracks.size() - 1;
}

extern "C" int cem_board_add_via(cem_board *b, double x, double y, double drill, double size, const char *layer_a,
                                 const char *layer_b, int net) {
    if (!b || !layer_a || !layer_b) return -1;
    b->vias.push_back(Via{Pt{x, y}, drill, size, layer_a, layer_b, net});
    return (int)b->vias.size() - 1;
}

extern "C" int cem_board_add_poly(cem_board *b, int npts, const double *xy, const char *layer, int net) {
    if (!b || !layer || npts < 0 || (npts && !xy)) return -1;
    Poly q;
    for (int i = 0; i < npts; i++) q.pts.push_back(Pt{xy[2 * i], xy[2 * i + 1]});
    q.layer = layer;
    q.net = net;
    b->polys.push_back(q);
    return (int)b->polys.size() - 1;
}

extern "C" int cem_board_set_net(cem_board *b, int net, const char *name) {
    if (!b || !name) return -1;
    b->set_net(net, name);
    return 0;
}

extern "C" int cem_board_set_outline(cem_board *b, double x0, double y0, double x1, double y1) {
    if (!b) return -1;
    b->omin = Pt{x0, y0};
    b->omax = Pt{x1, y1};
    return 0;
}

extern "C" int cem_board_count(const cem_board *b, int what) {
    if (!b) return -1;
    switch (what) {
    case 0: return (int)b->layers.size();
    case 1: return (int)b->tracks.size();
    case 2: return (int)b->vias.size();
    case 3: return (int)b->polys.size();
    case 4: return (int)b->nets.size();
    default: return -1;
    }
}

extern "C" int cem_board_layer(const cem_board *b, int i, char *name, int cap, int *type, double *v, int *mask) {
    if (!b || i < 0 || i >= (int)b->layers.size()) return -1;
    const Layer &l = b->layers[i];
    copy_name(l.name, name, cap);
    if (type) *type = l.type;
    if (mask) *mask = l.mask;
    if (v) {
        v[0] = l.thickness;
        v[1] = l.eps_r;
        v[2] = l.loss_tangent;
        v[3] = l.roughness;
        v[4] = l.z0;
        v[5] = l.z1;
    }
    return 0;
}

extern "C" int cem_board_track(const cem_board *b, int i, double *v, char *layer, int cap, int *net) {
    if (!b || i < 0 || i >= (int)b->tracks.size()) return -1;
    const Track &t = b->tracks[i];
    if (v) {
        v[0] = t.a.x;
        v[1] = t.a.y;
        v[2] = t.b.x;
        v[3] = t.b.y;
        v[4] = t.width;
    }
    copy_name(t.layer, layer, cap);
    if (net) *net = t.net;
    return 0;
}

extern "C" int cem_board_via(const cem_board *b, int i, double *v, char *layer_a, char *layer_b, int cap, int *net) {
    if (!b || i < 0 || i >= (int)b->vias.size()) return -1;
    const Via &w = b->vias[i];
    if (v) {
        v[0] = w.at.x;
        v[1] = w.at.y;
        v[2] = w.drill;
        v[3] = w.size;
    }
    copy_name(w.la, layer_a, cap);
    copy_name(w.lb, layer_b, cap);
    if (net) *net = w.net;
    return 0;
}

extern "C" int cem_board_poly(const cem_board *b, int i, double *xy, int xy_cap, char *layer, int cap, int *net) {
    if (!b || i < 0 || i >= (int)b->polys.size()) return -1;
    const Poly &q = b->polys[i];
    for (int k = 0; k < (int)q.pts.size() && k < xy_cap && xy; k++) {
        xy[2 * k] = q.pts[k].x;
        xy[2 * k + 1] = q.pts[k].y;
    }
    copy_name(q.layer, layer, cap);
    if (net) *net = q.net;
    return (int)q.pts.size();
}

extern "C" int cem_board_net(const cem_board *b, int i, int *number, char *name, int cap) {
    if (!b || i < 0 || i >= (int)b->nets.size()) return -1;
    if (number) *number = b->nets[i].first;
    copy_name(b->nets[i].second, name, cap);
    return 0;
}

extern "C" int cem_board_outline(const cem_board *b, double *v) {
    if (!b || !v) return -1;
    v[0] = b->omin.x;
    v[1] = b->omin.y;
    v[2] = b->omax.x;
    v[3] = b->omax.y;
    return 0;
}

extern "C" int cem_board_copper_at(const cem_board *b, const char *layer, double x, double y) {
    if (!b || !layer) return -1;
    return copper_at(*b, layer, Pt{x, y}) ? 1 : 0;
}

extern "C" int cem_board_copper_row(const cem_board *b, const char *layer, double y, int n, const double *x,
                                    unsigned char *out) {
    if (!b || !layer || n < 0 || (n && (!x || !out))) return -1;
    const Row r = copper_row(*b, layer, y);
    for (int i = 0; i < n; i++) out[i] = r.has(x[i]) ? 1 : 0;
    return 0;
}

extern "C" int cem_board_copper_edges(const cem_board *b, const char *layer, int axis, double *x, int *metal_hi,
                                      int cap) {
    if (!b || !layer || (axis != 0 && axis != 1)) return -1;
    const auto e = copper_edges(*b, layer, axis);
    for (int i = 0; i < (int)e.size() && i < cap; i++) {
        if (x) x[i] = e[i].x;
        if (metal_hi) metal_hi[i] = e[i].metal_hi;
    }
    return (int)e.size();
}

extern "C" double cem_lambda_over_20(double fmax, double er) { return lambda_over_20(fmax, er); }

extern "C" int cem_mesh_axis(const cem_axis_features *f, double hmin, double hmax, double ratio, double *nodes,
                             int cap, double *report, char *warnings, int wcap) {
    return guarded([&] {
        if (!f || !(hmin > 0) || !(hmax > 0) || !(f->hi > f->lo)) throw Fail("mesh axis: bad features or sizes");
        if ((f->nsheet > 0 && (!f->sheet_x || !f->sheet_metal_hi)) || (f->nplanes > 0 && !f->planes) ||
            (f->nspans > 0 && !f->spans) || (f->ncaps > 0 && !f->caps) || (f->nrefine > 0 && !f->refine) ||
            (f->nextra > 0 && !f->extra))
            throw Fail("mesh axis: a feature count without its array");
        for (int i = 0; i < f->nspans; i++)
            if (!(f->spans[3 * i + 2] >= 1)) throw Fail("mesh axis: a span needs at least one cell");
        Features F;
        F.lo = f->lo;
        F.hi = f->hi;
        for (int i = 0; i < f->nsheet; i++) F.sheets.push_back(Edge{f->sheet_x[i], f->sheet_metal_hi[i] ? 1 : 0});
        F.planes.assign(f->planes, f->planes + std::max(0, f->nplanes));
        for (int i = 0; i < f->nspans; i++)
            F.spans.push_back(Span{f->spans[3 * i], f->spans[3 * i + 1], (int)f->spans[3 * i + 2]});
        for (int i = 0; i < f->ncaps; i++) F.caps.push_back(Cap{f->caps[3 * i], f->caps[3 * i + 1], f->caps[3 * i + 2]});
        F.has_refine = f->nrefine >= 0;
        if (F.has_refine) F.refine.assign(f->refine, f->refine + f->nrefine);
        F.extra.assign(f->extra, f->extra + std::max(0, f->nextra));
        const AxisMesh m = mesh_axis(F, hmin, hmax, ratio);
        for (int i = 0; i < (int)m.nodes.size() && i < cap && nodes; i++) nodes[i] = m.nodes[i];
        if (report) {
            report[0] = (double)m.nodes.size() - 1;
            report[1] = m.min_cell;
            report[2] = m.max_cell;
            report[3] = m.max_ratio;
            report[4] = (double)m.warnings.size();
        }
        if (warnings && wcap > 0) {
            std::string all;
            for (auto &w : m.warnings) all += w + "\n";
            std::snprintf(warnings, (size_t)wcap, "%s", all.c_str());
        }
        return (int)m.nodes.size();
    });
}

extern "C" int cem_mesh_with_pml(const double *nodes, int n, int npml, double h, int lo, int hi, double *out) {
    if (!nodes || n < 1 || npml < 0 || !out) return -1;
    const auto v = with_pml(std::vector<double>(nodes, nodes + n), npml, h, lo != 0, hi != 0);
    std::copy(v.begin(), v.end(), out);
    return (int)v.size();
}

// pre-run estimate (mesher.ts:139-156): the bytes of the reference's own GPU and
// CPU storage models
extern "C" int cem_mesh_estimate(const double *x, int nx1, const double *y, int ny1, const double *z, int nz1,
                                 double t_sim, int gpu, int npml, double *out) {
    if (!x || !y || !z || nx1 < 2 || ny1 < 2 || nz1 < 2 || !out) return -1;
    const double n[3] = {(double)(nx1 - 1), (double)(ny1 - 1), (double)(nz1 - 1)};
    const double N = (n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    auto min_cell = [](const double *a, int m) {
        double v = kInf;
        for (int i = 0; i + 1 < m; i++) v = std::min(v, a[i + 1] - a[i]);
        return v;
    };
    const double mx = min_cell(x, nx1), my = min_cell(y, ny1), mz = min_cell(z, nz1);
    const double dt = 0.99 / (299792458.0 * std::sqrt(1 / (mx * mx) + 1 / (my * my) + 1 / (mz * mz)));
    const double steps = std::ceil(t_sim / dt);
    const double np = npml < 0 ? 8 : npml;
    const double per[3] = {(n[1] + 1) * (n[2] + 1), (n[0] + 1) * (n[2] + 1), (n[0] + 1) * (n[1] + 1)};
    double psi = 0;
    for (double q : per) psi += 4 * 2 * np * q;
    const double bytes = gpu ? 4 * (N * 12 + psi) : N * 8 * (12 + 2) + N + 8 * psi;
    out[0] = n[0] * n[1] * n[2];
    out[1] = N;
    out[2] = dt;
    out[3] = steps;
    out[4] = bytes;
    out[5] = n[0] * n[1] * n[2] * steps;
    return 0;
}

extern "C" int cem_board_plan(const cem_board *b, const cem_board_options *o, const cem_board_port *ports, int nports,
                              int *cells, double *x, double *y, double *z) {
    return guarded([&] {
        if (!b || nports < 0 || (nports && !ports)) throw Fail("bad board plan arguments");
        const auto p = plan_board(*b, options_of(o), ports_of(ports, nports));
        if (cells) {
            cells[0] = p->nx;
            cells[1] = p->ny;
            cells[2] = p->nz;
        }
        if (x) std::copy(p->x.begin(), p->x.end(), x);
        if (y) std::copy(p->y.begin(), p->y.end(), y);
        if (z) std::copy(p->z.begin(), p->z.end(), z);
        return 0;
    });
}

extern "C" cem_sim *cem_board_build(const cem_board *b, const cem_board_options *o, const cem_board_port *ports,
                                    int nports, int *port_ids, double *t_stop) {
    cem_sim *sim = nullptr;
    try {
        if (!b || nports < 1 || !ports || !port_ids) throw Fail("a board run needs at least one port");
        const auto p = plan_board(*b, options_of(o), ports_of(ports, nports));
        sim = cem_create(p->nx, p->ny, p->nz, p->x.data(), p->y.data(), p->z.data());
        if (!sim) throw Fail(std::string("create: ") + cem_last_error());
        build_into(sim, *p, port_ids, t_stop);
        return sim;
    } catch (const std::exception &e) {
        const std::string msg = e.what();
        if (sim) cem_destroy(sim);
        cem::set_error(msg.c_str());
        return nullptr;
    }
}
