// Host geometry and initial states (CPU): neighbour tables in the reference
// numbering, node Jacobians (analytic or 6th-order compact), node metrics,
// half-point metrics, and the free-stream / vortex / Taylor-Green states.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace {

inline int64_t nid(int32_t i, int32_t j, int32_t k, int32_t ni, int32_t nj) { return (int64_t)i + (int64_t)ni * (j + (int64_t)nj * k); }
inline int32_t wrap(int32_t i, int32_t n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

}  // namespace

extern "C" void cfdx_params_default(cfdx_params *p)
{
    if (!p) return;
    p->mach = 1.5;      // Viscous.cl:30
    p->reyl = 500.0;    // Viscous.cl:31
    p->pr = 0.72;       // Viscous.cl:32
    p->free_t = 273.0;  // Viscous.cl:33
    p->c_mts = 0.03;    // Viscous.cl:490
    p->c_t = 10.0;      // Viscous.cl:491
    p->eps = 10e-8;     // Weno.cl:59
    p->precision = CFDX_FP64;
    p->viscous = 1;
    p->sgs = CFDX_SGS_MTS;
    p->pad_ = 0;
    p->c_s = 0.17;
}

extern "C" int cfdx_geom_periodic_table(int32_t ni, int32_t nj, int32_t nk, const double *xyz, cfdx_grid *G)
{
    if (ni < 1 || nj < 1 || nk < 1 || !G) return cfdx::fail(CFDX_E_INVALID, "periodic_table: bad size or null table");
    std::memset(&G[0], 0, sizeof(cfdx_grid));  // dummy node 0: all neighbours 0, loc 0
    for (int32_t k = 0; k < nk; ++k)
        for (int32_t j = 0; j < nj; ++j)
            for (int32_t i = 0; i < ni; ++i) {
                const int64_t n = nid(i, j, k, ni, nj);
                cfdx_grid &g = G[n + 1];
                g.x = xyz ? xyz[3 * n] : 0.0;
                g.y = xyz ? xyz[3 * n + 1] : 0.0;
                g.z = xyz ? xyz[3 * n + 2] : 0.0;
                // Main.cpp:5675-5680 / U_Q.cl:51-56: N0 +j, N1 +i, N2 -j, N3 -i, N4 +k, N5 -k
                g.N0 = (int32_t)(1 + nid(i, wrap(j + 1, nj), k, ni, nj));
                g.N1 = (int32_t)(1 + nid(wrap(i + 1, ni), j, k, ni, nj));
                g.N2 = (int32_t)(1 + nid(i, wrap(j - 1, nj), k, ni, nj));
                g.N3 = (int32_t)(1 + nid(wrap(i - 1, ni), j, k, ni, nj));
                g.N4 = (int32_t)(1 + nid(i, j, wrap(k + 1, nk), ni, nj));
                g.N5 = (int32_t)(1 + nid(i, j, wrap(k - 1, nk), ni, nj));
                g.ID = 0;
                g.loc = 0;
            }
    return CFDX_OK;
}

extern "C" int cfdx_geom_coords_box(int32_t ni, int32_t nj, int32_t nk, const double L[3], double *xyz)
{
    if (ni < 1 || nj < 1 || nk < 1 || !L || !xyz) return cfdx::fail(CFDX_E_INVALID, "coords_box: bad argument");
    for (int32_t k = 0; k < nk; ++k)
        for (int32_t j = 0; j < nj; ++j)
            for (int32_t i = 0; i < ni; ++i) {
                const int64_t n = nid(i, j, k, ni, nj);
                xyz[3 * n] = i * (L[0] / ni);
                xyz[3 * n + 1] = j * (L[1] / nj);
                xyz[3 * n + 2] = k * (L[2] / nk);
            }
    return CFDX_OK;
}

extern "C" int cfdx_geom_coords_sinusoid(int32_t ni, int32_t nj, int32_t nk, const double L[3], double a,
                                         const double ph[3], double *xyz)
{
    if (ni < 1 || nj < 1 || nk < 1 || !L || !ph || !xyz) return cfdx::fail(CFDX_E_INVALID, "coords_sinusoid: bad argument");
    const double w0 = 2.0 * M_PI / L[0], w1 = 2.0 * M_PI / L[1], w2 = 2.0 * M_PI / L[2];
    for (int32_t k = 0; k < nk; ++k)
        for (int32_t j = 0; j < nj; ++j)
            for (int32_t i = 0; i < ni; ++i) {
                const int64_t n = nid(i, j, k, ni, nj);
                const double s1 = i * (L[0] / ni), s2 = j * (L[1] / nj), s3 = k * (L[2] / nk);
                xyz[3 * n] = s1 + a * std::sin(w1 * s2 + ph[0]) * std::sin(w2 * s3 + ph[0]);
                xyz[3 * n + 1] = s2 + a * std::sin(w2 * s3 + ph[1]) * std::sin(w0 * s1 + ph[1]);
                xyz[3 * n + 2] = s3 + a * std::sin(w0 * s1 + ph[2]) * std::sin(w1 * s2 + ph[2]);
            }
    return CFDX_OK;
}

extern "C" int cfdx_geom_jacobian_sinusoid(int32_t ni, int32_t nj, int32_t nk, const double L[3], double a,
                                           const double ph[3], double *jac)
{
    if (ni < 1 || nj < 1 || nk < 1 || !L || !ph || !jac) return cfdx::fail(CFDX_E_INVALID, "jacobian_sinusoid: bad argument");
    const double w0 = 2.0 * M_PI / L[0], w1 = 2.0 * M_PI / L[1], w2 = 2.0 * M_PI / L[2];
    const double h0 = L[0] / ni, h1 = L[1] / nj, h2 = L[2] / nk;  // d s / d index
    for (int32_t k = 0; k < nk; ++k)
        for (int32_t j = 0; j < nj; ++j)
            for (int32_t i = 0; i < ni; ++i) {
                const int64_t n = nid(i, j, k, ni, nj);
                const double s1 = i * h0, s2 = j * h1, s3 = k * h2;
                double *J = jac + 9 * n;  // x_zeta x_eta x_xi y_zeta y_eta y_xi z_zeta z_eta z_xi
                J[0] = h0;
                J[1] = h1 * a * w1 * std::cos(w1 * s2 + ph[0]) * std::sin(w2 * s3 + ph[0]);
                J[2] = h2 * a * w2 * std::sin(w1 * s2 + ph[0]) * std::cos(w2 * s3 + ph[0]);
                J[3] = h0 * a * w0 * std::sin(w2 * s3 + ph[1]) * std::cos(w0 * s1 + ph[1]);
                J[4] = h1;
                J[5] = h2 * a * w2 * std::cos(w2 * s3 + ph[1]) * std::sin(w0 * s1 + ph[1]);
                J[6] = h0 * a * w0 * std::cos(w0 * s1 + ph[2]) * std::sin(w1 * s2 + ph[2]);
                J[7] = h1 * a * w1 * std::sin(w0 * s1 + ph[2]) * std::cos(w1 * s2 + ph[2]);
                J[8] = h2;
            }
    return CFDX_OK;
}

namespace {

// Solves the cyclic system (1/3) f'_{i-1} + f'_i + (1/3) f'_{i+1} = r_i (i mod n) in place,
// by the Sherman-Morrison split of the periodic corner terms.
void cyclic_solve(std::vector<double> &r)
{
    const int n = (int)r.size();
    const double a = 1.0 / 3.0, b = 1.0, c = 1.0 / 3.0;
    if (n == 1) { r[0] /= (a + b + c); return; }
    if (n == 2) {  // [[b, a+c], [a+c, b]]
        const double d0 = b, o = a + c, det = d0 * d0 - o * o;
        const double x0 = (d0 * r[0] - o * r[1]) / det, x1 = (d0 * r[1] - o * r[0]) / det;
        r[0] = x0; r[1] = x1; return;
    }
    const double gamma = -b;
    std::vector<double> diag(n, b), u(n, 0.0), cp(n), x(n), z(n);
    diag[0] = b - gamma;
    diag[n - 1] = b - a * c / gamma;
    u[0] = gamma; u[n - 1] = c;  // corner vector
    auto thomas = [&](const std::vector<double> &rhs, std::vector<double> &out) {
        cp[0] = c / diag[0];
        out[0] = rhs[0] / diag[0];
        for (int i = 1; i < n; ++i) {
            const double m = diag[i] - a * cp[i - 1];
            cp[i] = c / m;
            out[i] = (rhs[i] - a * out[i - 1]) / m;
        }
        for (int i = n - 2; i >= 0; --i) out[i] -= cp[i] * out[i + 1];
    };
    thomas(r, x);
    thomas(u, z);
    const double fact = (x[0] + a * x[n - 1] / gamma) / (1.0 + z[0] + a * z[n - 1] / gamma);
    for (int i = 0; i < n; ++i) r[i] = x[i] - fact * z[i];
}

}  // namespace

extern "C" int cfdx_geom_jacobian_compact_periodic(int32_t ni, int32_t nj, int32_t nk, const double L[3],
                                                   const double *xyz, double *jac)
{
    if (ni < 1 || nj < 1 || nk < 1 || !L || !xyz || !jac) return cfdx::fail(CFDX_E_INVALID, "jacobian_compact: bad argument");
    const int32_t dims[3] = {ni, nj, nk};
    for (int d = 0; d < 3; ++d) {
        const int32_t n = dims[d];
        const double h = L[d] / n;  // the coordinate along d grows by one period over the line
        std::vector<double> f(n), r(n);
        const int32_t o1 = d == 0 ? nj : ni, o2 = d == 2 ? nj : nk;
        for (int32_t b = 0; b < o2; ++b)
            for (int32_t a = 0; a < o1; ++a)
                for (int c = 0; c < 3; ++c) {
                    auto at = [&](int32_t s) -> int64_t {
                        int32_t i = d == 0 ? s : a, j = d == 1 ? s : (d == 0 ? a : b), k = d == 2 ? s : b;
                        return nid(i, j, k, ni, nj);
                    };
                    for (int32_t s = 0; s < n; ++s) f[s] = xyz[3 * at(s) + c] - (c == d ? s * h : 0.0);
                    // Grid_Transformation.cpp:45-53 interior right-hand side (del_zeta = 1)
                    for (int32_t s = 0; s < n; ++s)
                        r[s] = (-1.0 / 36.0) * f[wrap(s - 2, n)] - (7.0 / 9.0) * f[wrap(s - 1, n)]
                               + (7.0 / 9.0) * f[wrap(s + 1, n)] + (1.0 / 36.0) * f[wrap(s + 2, n)];
                    cyclic_solve(r);
                    for (int32_t s = 0; s < n; ++s) jac[9 * at(s) + 3 * c + d] = r[s] + (c == d ? h : 0.0);
                }
    }
    return CFDX_OK;
}

extern "C" int cfdx_geom_node_metrics(int64_t n, const double *jac, double *det, cfdx_metric *metric)
{
    if (n < 0 || !jac || !det || !metric) return cfdx::fail(CFDX_E_INVALID, "node_metrics: bad argument");
    for (int64_t i = 0; i < n; ++i) {
        const double *J = jac + 9 * i;
        const double x_zeta = J[0], x_eta = J[1], x_xi = J[2];
        const double y_zeta = J[3], y_eta = J[4], y_xi = J[5];
        const double z_zeta = J[6], z_eta = J[7], z_xi = J[8];
        // Grid_Transformation.cpp:447
        const double D = (((x_zeta) * ((y_eta * z_xi) - (y_xi * z_eta))) + ((x_eta) * ((y_xi * z_zeta) - (y_zeta * z_xi)))
                          + ((x_xi) * ((y_zeta * z_eta) - (y_eta * z_zeta))));
        det[i] = D;
        double *m = metric[i].node;
        // Grid_Transformation.cpp:1029-1037
        m[0] = (1.0 / D) * (y_eta * z_xi - y_xi * z_eta);
        m[1] = (1.0 / D) * (z_eta * x_xi - z_xi * x_eta);
        m[2] = (1.0 / D) * (x_eta * y_xi - x_xi * y_eta);
        m[3] = (1.0 / D) * (y_xi * z_zeta - y_zeta * z_xi);
        m[4] = (1.0 / D) * (z_xi * x_zeta - z_zeta * x_xi);
        m[5] = (1.0 / D) * (x_xi * y_zeta - x_zeta * y_xi);
        m[6] = (1.0 / D) * (y_zeta * z_eta - y_eta * z_zeta);
        m[7] = (1.0 / D) * (z_zeta * x_eta - z_eta * x_zeta);
        m[8] = (1.0 / D) * (x_zeta * y_eta - x_eta * y_zeta);
    }
    return CFDX_OK;
}

extern "C" int cfdx_geom_half_metrics(const cfdx_grid *G, int64_t n_sd, cfdx_metric *metric, const double *det,
                                      cfdx_deter *deter)
{
    if (!G || n_sd < 0 || !metric || !det || !deter) return cfdx::fail(CFDX_E_INVALID, "half_metrics: bad argument");
    // Main.cpp:2368-2435: m_{+1/2} = (3 m_{-2} - 25 m_{-1} + 150 m_0 + 150 m_{+1} - 25 m_{+2} + 3 m_{+3}) / 256
    // face slot order ip, im, jp, jm, kp, km; plus uses N1/N0/N4, minus N3/N2/N5
    for (int64_t i = 1; i <= n_sd; ++i) {
        for (int d = 0; d < 3; ++d) {
            auto plus = [&](int64_t g) -> int64_t { return d == 0 ? G[g].N1 : d == 1 ? G[g].N0 : G[g].N4; };
            auto minus = [&](int64_t g) -> int64_t { return d == 0 ? G[g].N3 : d == 1 ? G[g].N2 : G[g].N5; };
            const int64_t m1 = minus(i), m2 = minus(m1), m3 = minus(m2);
            const int64_t p1 = plus(i), p2 = plus(p1), p3 = plus(p2);
            const int64_t sp[6] = {m2, m1, i, p1, p2, p3};   // plus face stencil
            const int64_t sm[6] = {m3, m2, m1, i, p1, p2};   // minus face stencil
            double *fp = metric[i].face[2 * d];
            double *fm = metric[i].face[2 * d + 1];
            for (int c = 0; c < 9; ++c) {
                fp[c] = (1.0 / 256.0) * (3.0 * (metric[sp[0]].node[c]) - 25.0 * (metric[sp[1]].node[c]) + 150.0 * (metric[sp[2]].node[c])
                                         + 150.0 * (metric[sp[3]].node[c]) - 25.0 * (metric[sp[4]].node[c]) + 3.0 * (metric[sp[5]].node[c]));
                fm[c] = (1.0 / 256.0) * (3.0 * (metric[sm[0]].node[c]) - 25.0 * (metric[sm[1]].node[c]) + 150.0 * (metric[sm[2]].node[c])
                                         + 150.0 * (metric[sm[3]].node[c]) - 25.0 * (metric[sm[4]].node[c]) + 3.0 * (metric[sm[5]].node[c]));
            }
            const double dp = (1.0 / 256.0) * (3.0 * (det[sp[0]]) - 25.0 * (det[sp[1]]) + 150.0 * (det[sp[2]]) + 150.0 * (det[sp[3]])
                                               - 25.0 * (det[sp[4]]) + 3.0 * (det[sp[5]]));
            const double dm = (1.0 / 256.0) * (3.0 * (det[sm[0]]) - 25.0 * (det[sm[1]]) + 150.0 * (det[sm[2]]) + 150.0 * (det[sm[3]])
                                               - 25.0 * (det[sm[4]]) + 3.0 * (det[sm[5]]));
            double *dd = &deter[i].ip;
            dd[2 * d] = dp;
            dd[2 * d + 1] = dm;
        }
    }
    return CFDX_OK;
}

namespace {
void set_flow(cfdx_flow &f, double rho, double u, double v, double w, double p, double mach)
{
    f.rho = rho; f.u = u; f.v = v; f.w = w; f.p = p;
    f.t = (1.4 * mach * mach) * (p / rho);
    f.e = p / (0.4 * rho);
    f.a = std::sqrt(1.4 * p / rho);
    f.mu = 0.0;
}
}  // namespace

extern "C" int cfdx_geom_init_freestream(int64_t n, double mach, cfdx_flow *flow)
{
    if (n < 0 || !flow || !(mach > 0)) return cfdx::fail(CFDX_E_INVALID, "init_freestream: bad argument");
    // Flow_Initialisation.cpp:35-48
    const double p = 1.0 / (1.4 * mach * mach);
    for (int64_t i = 0; i <= n; ++i) set_flow(flow[i], (1.4 * mach * mach) * (p / 1.0), 1.0, 0.0, 0.0, p, mach);
    return CFDX_OK;
}

extern "C" int cfdx_geom_init_vortex(int64_t n, const double *xyz, double mach, double beta, double xc, double yc,
                                     cfdx_flow *flow)
{
    if (n < 0 || !xyz || !flow || !(mach > 0)) return cfdx::fail(CFDX_E_INVALID, "init_vortex: bad argument");
    const double g = 1.4;
    cfdx_geom_init_freestream(0, mach, flow);
    for (int64_t i = 0; i < n; ++i) {
        const double dx = xyz[3 * i] - xc, dy = xyz[3 * i + 1] - yc, r2 = dx * dx + dy * dy;
        // velocities in free-stream sound-speed units (c_inf = 1/M in free-stream velocity units)
        const double du = -(beta / (2.0 * M_PI)) * dy * std::exp(0.5 * (1.0 - r2)) / mach;
        const double dv = (beta / (2.0 * M_PI)) * dx * std::exp(0.5 * (1.0 - r2)) / mach;
        const double T = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * M_PI * M_PI) * std::exp(1.0 - r2);
        const double rho = std::pow(T, 1.0 / (g - 1.0));
        const double p = rho * T / (g * mach * mach);
        set_flow(flow[i + 1], rho, 1.0 + du, dv, 0.0, p, mach);
    }
    return CFDX_OK;
}

extern "C" int cfdx_geom_init_tgv(int64_t n, const double *xyz, double mach, cfdx_flow *flow)
{
    if (n < 0 || !xyz || !flow || !(mach > 0)) return cfdx::fail(CFDX_E_INVALID, "init_tgv: bad argument");
    cfdx_geom_init_freestream(0, mach, flow);
    const double pinf = 1.0 / (1.4 * mach * mach);
    for (int64_t i = 0; i < n; ++i) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const double u = std::sin(x) * std::cos(y) * std::cos(z);
        const double v = -std::cos(x) * std::sin(y) * std::cos(z);
        const double p = pinf + (1.0 / 16.0) * (std::cos(2.0 * x) + std::cos(2.0 * y)) * (std::cos(2.0 * z) + 2.0);
        const double rho = 1.4 * mach * mach * p;  // T = 1
        set_flow(flow[i + 1], rho, u, v, 0.0, p, mach);
    }
    return CFDX_OK;
}
