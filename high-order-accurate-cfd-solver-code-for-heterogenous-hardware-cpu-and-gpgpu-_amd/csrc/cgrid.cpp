// Host geometry of bounded blocks: an algebraic C-grid around a NACA 4-digit section extruded
// in span, a sinusoidally mapped box, and the metrics of any stored (padded) block by the compact
// 6th-order scheme of Grid_Transformation.cpp:45-153 with its boundary closures (alpha = 5 at the
// end node, 2/11 at the next, :55-72), node metrics (:447, :1029-1037) and half-point metrics
// (Main.cpp:2368-2435). Lines through a C-grid wake cut are solved as one line.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace {

struct Pad {
    int n[3], off[3], dims[3];
    bool periodic[3];
    size_t at(int i, int j, int k) const   // owned coordinates, possibly in the ghost layers
    {
        return (size_t)(i + off[0]) + (size_t)dims[0] * ((size_t)(j + off[1]) + (size_t)dims[1] * (size_t)(k + off[2]));
    }
    size_t count() const { return (size_t)dims[0] * dims[1] * dims[2]; }
};

Pad pad_of(const cfdx_block_desc *d)
{
    Pad p;
    const int n[3] = {d->ni, d->nj, d->nk};
    for (int q = 0; q < 3; ++q) {
        p.n[q] = n[q];
        p.periodic[q] = (d->periodic >> q) & 1;
        p.off[q] = p.periodic[q] ? 0 : 3;
        p.dims[q] = n[q] + 2 * p.off[q];
    }
    return p;
}

bool cut_of(const cfdx_block_desc *d) { return !(d->periodic & 2) && d->bc[2] == CFDX_BC_WALL_CUT; }

// d/ds of f over a non-periodic line (unit spacing): compact interior scheme and the boundary
// closures; Thomas algorithm
void compact_line(const std::vector<double> &f, std::vector<double> &df)
{
    const int n = (int)f.size();
    std::vector<double> a(n, 0.0), b(n, 1.0), c(n, 0.0), r(n, 0.0);
    for (int s = 0; s < n; ++s) {
        if (s == 0) {
            c[s] = 5.0;
            r[s] = (-197.0 / 60.0) * f[0] - (5.0 / 12.0) * f[1] + 5.0 * f[2] - (5.0 / 3.0) * f[3] + (5.0 / 12.0) * f[4] - (1.0 / 20.0) * f[5];
        } else if (s == n - 1) {
            a[s] = 5.0;
            r[s] = (197.0 / 60.0) * f[n - 1] + (5.0 / 12.0) * f[n - 2] - 5.0 * f[n - 3] + (5.0 / 3.0) * f[n - 4] - (5.0 / 12.0) * f[n - 5]
                   + (1.0 / 20.0) * f[n - 6];
        } else if (s == 1) {
            a[s] = c[s] = 2.0 / 11.0;
            r[s] = (-20.0 / 33.0) * f[0] - (35.0 / 132.0) * f[1] + (34.0 / 33.0) * f[2] - (7.0 / 33.0) * f[3] + (2.0 / 33.0) * f[4]
                   - (1.0 / 132.0) * f[5];
        } else if (s == n - 2) {
            a[s] = c[s] = 2.0 / 11.0;
            r[s] = (20.0 / 33.0) * f[n - 1] + (35.0 / 132.0) * f[n - 2] - (34.0 / 33.0) * f[n - 3] + (7.0 / 33.0) * f[n - 4]
                   - (2.0 / 33.0) * f[n - 5] + (1.0 / 132.0) * f[n - 6];
        } else {
            a[s] = c[s] = 1.0 / 3.0;
            r[s] = (7.0 / 9.0) * (f[s + 1] - f[s - 1]) + (1.0 / 36.0) * (f[s + 2] - f[s - 2]);
        }
    }
    for (int s = 1; s < n; ++s) {
        const double w = a[s] / b[s - 1];
        b[s] -= w * c[s - 1];
        r[s] -= w * r[s - 1];
    }
    df.assign(n, 0.0);
    df[n - 1] = r[n - 1] / b[n - 1];
    for (int s = n - 2; s >= 0; --s) df[s] = (r[s] - c[s] * df[s + 1]) / b[s];
}

// periodic line: (1/3) f'_{s-1} + f'_s + (1/3) f'_{s+1} = interior right-hand side, cyclic, by
// the Sherman-Morrison split (period added to the coordinate along the line by the caller)
void compact_cyclic(const std::vector<double> &f, std::vector<double> &df)
{
    const int n = (int)f.size();
    auto w = [&](int s) { return ((s % n) + n) % n; };
    std::vector<double> r(n);
    for (int s = 0; s < n; ++s) r[s] = (7.0 / 9.0) * (f[w(s + 1)] - f[w(s - 1)]) + (1.0 / 36.0) * (f[w(s + 2)] - f[w(s - 2)]);
    const double a = 1.0 / 3.0, b = 1.0, c = 1.0 / 3.0, gamma = -b;
    std::vector<double> diag(n, b), u(n, 0.0), cp(n), x(n), z(n);
    diag[0] = b - gamma;
    diag[n - 1] = b - a * c / gamma;
    u[0] = gamma;
    u[n - 1] = c;
    auto thomas = [&](const std::vector<double> &rhs, std::vector<double> &out) {
        cp[0] = c / diag[0];
        out[0] = rhs[0] / diag[0];
        for (int s = 1; s < n; ++s) {
            const double m = diag[s] - a * cp[s - 1];
            cp[s] = c / m;
            out[s] = (rhs[s] - a * out[s - 1]) / m;
        }
        for (int s = n - 2; s >= 0; --s) out[s] -= cp[s] * out[s + 1];
    };
    thomas(r, x);
    thomas(u, z);
    const double fact = (x[0] + a * x[n - 1] / gamma) / (1.0 + z[0] + a * z[n - 1] / gamma);
    df.resize(n);
    for (int s = 0; s < n; ++s) df[s] = x[s] - fact * z[s];
}

// NACA 4-digit section (closed trailing edge): point at chordwise x on the upper or lower surface
void naca_point(double m, double p, double t, double x, bool upper, double &X, double &Y)
{
    const double yt = 5.0 * t * (0.2969 * std::sqrt(std::max(x, 0.0)) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
    double yc = 0.0, dyc = 0.0;
    if (m > 0.0) {
        if (x < p) { yc = m / (p * p) * (2.0 * p * x - x * x); dyc = 2.0 * m / (p * p) * (p - x); }
        else { yc = m / ((1 - p) * (1 - p)) * ((1.0 - 2.0 * p) + 2.0 * p * x - x * x); dyc = 2.0 * m / ((1 - p) * (1 - p)) * (p - x); }
    }
    const double th = std::atan(dyc);
    if (upper) { X = x - yt * std::sin(th); Y = yc + yt * std::cos(th); }
    else { X = x + yt * std::sin(th); Y = yc - yt * std::cos(th); }
}

}  // namespace

extern "C" int cfdx_geom_cgrid_desc(int32_t ni, int32_t nj, int32_t nk, int32_t precision, const cfdx_cgrid *g, double aoa,
                                    cfdx_block_desc *out)
{
    if (!g || !out) return cfdx::fail(CFDX_E_INVALID, "cgrid_desc: null argument");
    if (g->n_wake < 2 || ni - 2 * g->n_wake < 8) return cfdx::fail(CFDX_E_INVALID, "cgrid_desc: need n_wake >= 2 and >= 8 airfoil points");
    std::memset(out, 0, sizeof *out);
    out->ni = ni; out->nj = nj; out->nk = nk;
    out->periodic = 4;
    out->precision = precision;
    out->bc[0] = CFDX_BC_OUTFLOW;
    out->bc[1] = CFDX_BC_OUTFLOW;
    out->bc[2] = CFDX_BC_WALL_CUT;
    out->bc[3] = CFDX_BC_FARFIELD;
    out->cut_i0 = g->n_wake;
    out->cut_i1 = ni - 1 - g->n_wake;
    out->aoa = aoa;
    return CFDX_OK;
}

extern "C" int cfdx_geom_cgrid_coords(const cfdx_block_desc *d, const cfdx_cgrid *g, double *xyz)
{
    if (!d || !g || !xyz) return cfdx::fail(CFDX_E_INVALID, "cgrid_coords: null argument");
    if (d->periodic != 4 || !cut_of(d)) return cfdx::fail(CFDX_E_INVALID, "cgrid_coords: needs a C-grid descriptor (cfdx_geom_cgrid_desc)");
    if (!(g->t > 0 && g->radius > 1 && g->wake > 0 && g->wall > 0 && g->span > 0) || (g->m > 0 && !(g->p > 0 && g->p < 1)))
        return cfdx::fail(CFDX_E_INVALID, "cgrid_coords: bad section or grid parameters");
    const Pad P = pad_of(d);
    const int ni = d->ni, nj = d->nj, nk = d->nk, nw = g->n_wake, naf = ni - 2 * nw;
    // airfoil i = nw .. ni-1-nw: lower surface from the trailing edge to the leading edge, then the
    // upper surface back, cosine-clustered at both edges
    auto airfoil = [&](int i, double &X, double &Y, double &tpar) {
        tpar = (double)(i - nw) / (double)(naf - 1);
        const double xc = 0.5 * (1.0 + std::cos(2.0 * M_PI * tpar));
        naca_point(g->m, g->p, g->t, xc, tpar > 0.5, X, Y);
    };
    // wake spacing grows geometrically from the trailing-edge spacing of the section
    double X1, Y1, tp;
    airfoil(nw + 1, X1, Y1, tp);
    const double d0 = std::max(std::hypot(X1 - 1.0, Y1), 1e-6);
    double r = 1.0001;
    {
        double lo = 1.0 + 1e-9, hi = 50.0;
        auto len = [&](double q) { return d0 * (std::pow(q, nw) - 1.0) / (q - 1.0); };
        if (len(hi) < g->wake) return cfdx::fail(CFDX_E_INVALID, "cgrid_coords: wake too long for %d wake points", nw);
        if (len(lo) > g->wake) r = lo;
        else {
            for (int it = 0; it < 200; ++it) {
                const double mid = 0.5 * (lo + hi);
                if (len(mid) < g->wake) lo = mid; else hi = mid;
            }
            r = 0.5 * (lo + hi);
        }
    }
    auto wake_x = [&](int s) { return 1.0 + d0 * (std::pow(r, s) - 1.0) / (r - 1.0); };   // s steps behind the edge
    // wall-normal stretching f(j) = sinh(beta j / (nj-1)) / sinh(beta), odd in j (ghost rows below
    // the wall and the cut are mirror images), first spacing f(1) = wall / radius
    double beta;
    {
        const double want = g->wall / g->radius;
        double lo = 1e-6, hi = 60.0;
        auto f1 = [&](double b) { return std::sinh(b / (nj - 1)) / std::sinh(b); };
        if (f1(lo) < want) return cfdx::fail(CFDX_E_INVALID, "cgrid_coords: wall spacing too large for %d normal points", nj);
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (f1(mid) > want) lo = mid; else hi = mid;
        }
        beta = 0.5 * (lo + hi);
    }
    auto fj = [&](int j) { return std::sinh(beta * j / (nj - 1)) / std::sinh(beta); };
    // Far field: the wake columns end on the lines y = -R / +R, uniformly in x over the wake length;
    // the airfoil columns end on a semicircle of radius R around the trailing edge, its angles
    // clustered at both ends so that the far-field spacing is continuous at the junction with the
    // wake. Airfoil columns are straight lines from the section to their far-field point (odd in j
    // through the wall). Wake columns lean from the cut toward their far-field point with the even
    // weight h(j) = f^2 / sqrt(f^2 + c^2) (normalised to 1 at the far field), c = d0 / R: they leave
    // the cut normal to it over one trailing-edge spacing and lean like the airfoil columns beyond,
    // and a ghost row below the cut is exactly the row across it. Beyond the far field a wake column
    // continues by (h - 1) D(s), D the running maximum over s of the lean xf - x0: the same line near
    // the edge, and never converging where the cut spacing outgrows the far-field spacing.
    const double cw = d0 / g->radius, hnorm = std::sqrt(1.0 + cw * cw);
    auto hj = [&](int j) {
        const double e = fj(j);
        return e * e / std::sqrt(e * e + cw * cw) * hnorm;
    };
    std::vector<double> lean(nw + P.off[0] + 1);
    for (int s = 0; s < (int)lean.size(); ++s) {
        const double d = 1.0 + g->wake * (double)s / (double)nw - wake_x(s);
        lean[s] = s ? std::max(lean[s - 1], d) : d;
    }
    const double dth_end = (g->wake / nw) / g->radius;          // wanted end angle step
    double amod = dth_end * (naf - 1) / M_PI - 1.0;            // theta(t) = pi (t + a sin(2 pi t) / (2 pi))
    amod = std::min(0.8, std::max(-0.8, amod));
    auto theta = [&](int i) {
        const double t = (double)(i - nw) / (double)(naf - 1);
        return -0.5 * M_PI - M_PI * (t + amod * std::sin(2.0 * M_PI * t) / (2.0 * M_PI));
    };
    for (int k = -P.off[2]; k < nk + P.off[2]; ++k)
        for (int j = -P.off[1]; j < nj + P.off[1]; ++j)
            for (int i = -P.off[0]; i < ni + P.off[0]; ++i) {
                double X, Y;
                const double f = fj(j);
                if (i < nw || i > ni - 1 - nw) {
                    const bool lower = i < nw;
                    const int s = lower ? nw - i : i - (ni - 1 - nw);   // steps behind the edge
                    const double x0 = wake_x(s), side = lower ? -1.0 : 1.0;
                    const double xf = 1.0 + g->wake * (double)s / (double)nw;
                    X = j < nj ? x0 + hj(j) * (xf - x0) : xf + (hj(j) - 1.0) * lean[s];
                    Y = f * side * g->radius;
                } else {
                    double x0, y0;
                    airfoil(i, x0, y0, tp);
                    const double th = theta(i);
                    const double xf = 1.0 + g->radius * std::cos(th), yf = g->radius * std::sin(th);
                    X = x0 + f * (xf - x0);
                    Y = y0 + f * (yf - y0);
                }
                double *Xp = xyz + 3 * P.at(i, j, k);
                Xp[0] = X;
                Xp[1] = Y;
                Xp[2] = k * (g->span / nk);
            }
    return CFDX_OK;
}

extern "C" int cfdx_geom_box_coords(const cfdx_block_desc *d, const double L[3], double amp, const double ph[3], double *xyz)
{
    if (!d || !L || !ph || !xyz) return cfdx::fail(CFDX_E_INVALID, "box_coords: null argument");
    const Pad P = pad_of(d);
    const double h[3] = {L[0] / d->ni, L[1] / d->nj, L[2] / d->nk};
    const double w[3] = {2.0 * M_PI / L[0], 2.0 * M_PI / L[1], 2.0 * M_PI / L[2]};
    for (int k = -P.off[2]; k < d->nk + P.off[2]; ++k)
        for (int j = -P.off[1]; j < d->nj + P.off[1]; ++j)
            for (int i = -P.off[0]; i < d->ni + P.off[0]; ++i) {
                const double s1 = i * h[0], s2 = j * h[1], s3 = k * h[2];
                double *X = xyz + 3 * P.at(i, j, k);
                X[0] = s1 + amp * std::sin(w[1] * s2 + ph[0]) * std::sin(w[2] * s3 + ph[0]);
                X[1] = s2 + amp * std::sin(w[2] * s3 + ph[1]) * std::sin(w[0] * s1 + ph[1]);
                X[2] = s3 + amp * std::sin(w[0] * s1 + ph[2]) * std::sin(w[1] * s2 + ph[2]);
            }
    return CFDX_OK;
}

extern "C" int cfdx_geom_block_metrics(const cfdx_block_desc *d, const double *xyz, const double L[3], double *geo)
{
    if (!d || !xyz || !geo) return cfdx::fail(CFDX_E_INVALID, "block_metrics: null argument");
    const Pad P = pad_of(d);
    const size_t N = P.count();
    for (int q = 0; q < 3; ++q)
        if (P.periodic[q] && !L) return cfdx::fail(CFDX_E_INVALID, "block_metrics: the period L is needed for periodic directions");
    std::vector<double> jac(9 * N, 0.0);   // x_zeta x_eta x_xi y_zeta ... z_xi per stored node
    const bool cut = cut_of(d);
    const int ni = d->ni, nj = d->nj;
    auto wall_col = [&](int i) { return i >= d->cut_i0 && i <= d->cut_i1; };
    std::vector<double> f, df;
    for (int dir = 0; dir < 3; ++dir) {
        const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;   // the two other directions
        for (int b = -P.off[o2]; b < P.n[o2] + P.off[o2]; ++b)
            for (int a = -P.off[o1]; a < P.n[o1] + P.off[o1]; ++a) {
                auto idx = [&](int s) {
                    int x[3];
                    x[dir] = s; x[o1] = a; x[o2] = b;
                    return P.at(x[0], x[1], x[2]);
                };
                if (P.periodic[dir]) {
                    const int n = P.n[dir];
                    const double per = L[dir];
                    for (int c = 0; c < 3; ++c) {
                        f.assign(n, 0.0);
                        for (int s = 0; s < n; ++s) f[s] = xyz[3 * idx(s) + c] - (c == dir ? s * (per / n) : 0.0);
                        compact_cyclic(f, df);
                        for (int s = 0; s < n; ++s) jac[9 * idx(s) + 3 * c + dir] = df[s] + (c == dir ? per / n : 0.0);
                    }
                } else if (dir == 1 && cut && !wall_col(a)) {
                    // through the wake cut: the upper column (ni-1-i) from its top down to row 1, then the
                    // lower column (i) from row 0 up; each (column, row) pair once, from the lower side
                    if (a > ni - 1 - a) continue;
                    const int au = ni - 1 - a, top = nj + 2;
                    const int len = 2 * top + 1;
                    auto lidx = [&](int s, int col, int row) {
                        (void)s;
                        int x[3];
                        x[0] = col; x[1] = row; x[2] = b;
                        return P.at(x[0], x[1], x[2]);
                    };
                    for (int c = 0; c < 3; ++c) {
                        f.assign(len, 0.0);
                        for (int s = 0; s < len; ++s) {
                            const size_t g = s < top ? lidx(s, au, top - s) : lidx(s, a, s - top);
                            f[s] = xyz[3 * g + c];
                        }
                        compact_line(f, df);
                        // lower column row j <-> s = top + j (j from -3); upper column row j' <-> s = top - j'
                        for (int j = -P.off[1]; j < nj + P.off[1]; ++j) {
                            jac[9 * lidx(0, a, j) + 3 * c + 1] = df[top + j];
                            if (au != a) jac[9 * lidx(0, au, j) + 3 * c + 1] = -df[top - j];
                        }
                    }
                } else {
                    const int n = P.dims[dir], s0 = -P.off[dir];
                    for (int c = 0; c < 3; ++c) {
                        f.assign(n, 0.0);
                        for (int s = 0; s < n; ++s) f[s] = xyz[3 * idx(s0 + s) + c];
                        compact_line(f, df);
                        for (int s = 0; s < n; ++s) jac[9 * idx(s0 + s) + 3 * c + dir] = df[s];
                    }
                }
            }
    }
    // det and node metrics at every stored node
    double *det = geo, *m[9];
    for (int c = 0; c < 9; ++c) m[c] = geo + (1 + c) * N;
    for (size_t g = 0; g < N; ++g) {
        const double *J = &jac[9 * g];
        const double x_zeta = J[0], x_eta = J[1], x_xi = J[2], y_zeta = J[3], y_eta = J[4], y_xi = J[5], z_zeta = J[6],
                     z_eta = J[7], z_xi = J[8];
        const double D = (((x_zeta) * ((y_eta * z_xi) - (y_xi * z_eta))) + ((x_eta) * ((y_xi * z_zeta) - (y_zeta * z_xi)))
                          + ((x_xi) * ((y_zeta * z_eta) - (y_eta * z_zeta))));
        det[g] = D;
        const double iD = D != 0.0 ? 1.0 / D : 0.0;
        m[0][g] = iD * (y_eta * z_xi - y_xi * z_eta);
        m[1][g] = iD * (z_eta * x_xi - z_xi * x_eta);
        m[2][g] = iD * (x_eta * y_xi - x_xi * y_eta);
        m[3][g] = iD * (y_xi * z_zeta - y_zeta * z_xi);
        m[4][g] = iD * (z_xi * x_zeta - z_zeta * x_xi);
        m[5][g] = iD * (x_xi * y_zeta - x_zeta * y_xi);
        m[6][g] = iD * (y_zeta * z_eta - y_eta * z_zeta);
        m[7][g] = iD * (z_zeta * x_eta - z_eta * x_zeta);
        m[8][g] = iD * (x_zeta * y_eta - x_eta * y_zeta);
    }
    // +1/2-face metrics of the faces the sweeps compute: along-coordinate -1 .. n-1 in a bounded
    // direction, 0 .. n-1 (wrapped) in a periodic one, on owned lines across
    double *kf = geo + 10 * N, *dfc = geo + 19 * N;
    std::fill(kf, kf + 9 * N, 0.0);
    std::fill(dfc, dfc + 3 * N, 0.0);
    for (int dir = 0; dir < 3; ++dir) {
        const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
        const int n = P.n[dir], lo = P.periodic[dir] ? 0 : -1;
        for (int b = 0; b < P.n[o2]; ++b)
            for (int a = 0; a < P.n[o1]; ++a)
                for (int s = lo; s < n; ++s) {
                    size_t st[6];
                    for (int o = -2; o <= 3; ++o) {
                        int x[3];
                        int c = s + o;
                        if (P.periodic[dir]) c = ((c % n) + n) % n;
                        x[dir] = c; x[o1] = a; x[o2] = b;
                        st[o + 2] = P.at(x[0], x[1], x[2]);
                    }
                    int x[3];
                    x[dir] = s; x[o1] = a; x[o2] = b;
                    const size_t g = P.at(x[0], x[1], x[2]);
                    auto interp = [&](const double *v) {
                        return (1.0 / 256.0) * (3.0 * (v[st[0]]) - 25.0 * (v[st[1]]) + 150.0 * (v[st[2]]) + 150.0 * (v[st[3]])
                                                - 25.0 * (v[st[4]]) + 3.0 * (v[st[5]]));
                    };
                    for (int c = 0; c < 3; ++c) kf[(3 * dir + c) * N + g] = interp(m[3 * dir + c]);
                    dfc[dir * N + g] = interp(det);
                }
    }
    return CFDX_OK;
}
