// Device-side set-up of a periodic block: the geometry of the seeded sinusoidal map (exact
// Jacobian -> det and node metrics, Grid_Transformation.cpp:447, :1029-1037; half-point
// metrics, Main.cpp:2368-2435), the initial state (free stream, isentropic vortex or
// Taylor-Green; Flow_Initialisation.cpp:35-48 for the free stream) and the CFL time step.
// The same formulas as the host functions of geometry.cpp, so a request never builds its grid
// on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "block.hpp"

namespace cfdx {
namespace blk {
namespace {

struct MapArgs {
    double L[3], amp, ph[3];
};

__device__ __forceinline__ void node_coords(const MapArgs &a, int ni, int nj, int nk, int i, int j, int k, double x[3],
                                            double s[3])
{
    const double w0 = 2.0 * M_PI / a.L[0], w1 = 2.0 * M_PI / a.L[1], w2 = 2.0 * M_PI / a.L[2];
    s[0] = i * (a.L[0] / ni); s[1] = j * (a.L[1] / nj); s[2] = k * (a.L[2] / nk);
    x[0] = s[0] + a.amp * sin(w1 * s[1] + a.ph[0]) * sin(w2 * s[2] + a.ph[0]);
    x[1] = s[1] + a.amp * sin(w2 * s[2] + a.ph[1]) * sin(w0 * s[0] + a.ph[1]);
    x[2] = s[2] + a.amp * sin(w0 * s[0] + a.ph[2]) * sin(w1 * s[1] + a.ph[2]);
}

__global__ void k_geom(BlockDev B, MapArgs a)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= B.n) return;
    const int i = (int)(g % (uint32_t)B.ni), j = (int)((g / (uint32_t)B.ni) % (uint32_t)B.nj), k = (int)(g / ((uint32_t)B.ni * B.nj));
    const double w0 = 2.0 * M_PI / a.L[0], w1 = 2.0 * M_PI / a.L[1], w2 = 2.0 * M_PI / a.L[2];
    const double h0 = a.L[0] / B.ni, h1 = a.L[1] / B.nj, h2 = a.L[2] / B.nk;
    const double s1 = i * h0, s2 = j * h1, s3 = k * h2;
    const double x_zeta = h0;
    const double x_eta = h1 * a.amp * w1 * cos(w1 * s2 + a.ph[0]) * sin(w2 * s3 + a.ph[0]);
    const double x_xi = h2 * a.amp * w2 * sin(w1 * s2 + a.ph[0]) * cos(w2 * s3 + a.ph[0]);
    const double y_zeta = h0 * a.amp * w0 * sin(w2 * s3 + a.ph[1]) * cos(w0 * s1 + a.ph[1]);
    const double y_eta = h1;
    const double y_xi = h2 * a.amp * w2 * cos(w2 * s3 + a.ph[1]) * sin(w0 * s1 + a.ph[1]);
    const double z_zeta = h0 * a.amp * w0 * cos(w0 * s1 + a.ph[2]) * sin(w1 * s2 + a.ph[2]);
    const double z_eta = h1 * a.amp * w1 * sin(w0 * s1 + a.ph[2]) * cos(w1 * s2 + a.ph[2]);
    const double z_xi = h2;
    const double D = (((x_zeta) * ((y_eta * z_xi) - (y_xi * z_eta))) + ((x_eta) * ((y_xi * z_zeta) - (y_zeta * z_xi)))
                      + ((x_xi) * ((y_zeta * z_eta) - (y_eta * z_zeta))));
    const double iD = 1.0 / D;
    double *det = const_cast<double *>(B.det);
    double *m[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) m[c] = const_cast<double *>(B.m[c]);
    det[g] = D;
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

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// +1/2-face metrics of each direction from the node values along it (periodic wrap)
__global__ void k_half(BlockDev B)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= B.n) return;
    const int i = (int)(g % (uint32_t)B.ni), j = (int)((g / (uint32_t)B.ni) % (uint32_t)B.nj), k = (int)(g / ((uint32_t)B.ni * B.nj));
    for (int d = 0; d < 3; ++d) {
        uint32_t s[6];
        for (int o = -2; o <= 3; ++o) {
            int ii = i, jj = j, kk = k;
            if (d == 0) ii = wrap(i + o, B.ni);
            else if (d == 1) jj = wrap(j + o, B.nj);
            else kk = wrap(k + o, B.nk);
            s[o + 2] = (uint32_t)ii + (uint32_t)B.ni * ((uint32_t)jj + (uint32_t)B.nj * (uint32_t)kk);
        }
        auto interp = [&](const double *f) {
            return (1.0 / 256.0) * (3.0 * (f[s[0]]) - 25.0 * (f[s[1]]) + 150.0 * (f[s[2]]) + 150.0 * (f[s[3]]) - 25.0 * (f[s[4]])
                                    + 3.0 * (f[s[5]]));
        };
        for (int c = 0; c < 3; ++c) const_cast<double *>(B.kf[d][c])[g] = interp(B.m[3 * d + c]);
        const_cast<double *>(B.df[d])[g] = interp(B.det);
    }
}

struct InitArgs {
    MapArgs map;
    int kind;   // 0 free stream, 1 isentropic vortex, 2 Taylor-Green
    double mach, beta, xc, yc;
};

// primitives (rho, u, v, w, p) in double into Us, for k_base to turn into P and U0
__global__ void k_init(BlockDev B, InitArgs a)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= B.n) return;
    const int i = (int)(g % (uint32_t)B.ni), j = (int)((g / (uint32_t)B.ni) % (uint32_t)B.nj), k = (int)(g / ((uint32_t)B.ni * B.nj));
    double x[3], s[3];
    node_coords(a.map, B.ni, B.nj, B.nk, i, j, k, x, s);
    const double M = a.mach, pinf = 1.0 / (1.4 * M * M);
    double rho = (1.4 * M * M) * (pinf / 1.0), u = 1.0, v = 0.0, w = 0.0, p = pinf;
    if (a.kind == 1) {
        const double gm = 1.4;
        const double dx = x[0] - a.xc, dy = x[1] - a.yc, r2 = dx * dx + dy * dy;
        const double du = -(a.beta / (2.0 * M_PI)) * dy * exp(0.5 * (1.0 - r2)) / M;
        const double dv = (a.beta / (2.0 * M_PI)) * dx * exp(0.5 * (1.0 - r2)) / M;
        const double T = 1.0 - (gm - 1.0) * a.beta * a.beta / (8.0 * gm * M_PI * M_PI) * exp(1.0 - r2);
        rho = pow(T, 1.0 / (gm - 1.0));
        p = rho * T / (gm * M * M);
        u = 1.0 + du; v = dv; w = 0.0;
    } else if (a.kind == 2) {
        u = sin(x[0]) * cos(x[1]) * cos(x[2]);
        v = -cos(x[0]) * sin(x[1]) * cos(x[2]);
        w = 0.0;
        p = pinf + (1.0 / 16.0) * (cos(2.0 * x[0]) + cos(2.0 * x[1])) * (cos(2.0 * x[2]) + 2.0);
        rho = 1.4 * M * M * p;
    }
    B.Us[0][g] = rho; B.Us[1][g] = u; B.Us[2][g] = v; B.Us[3][g] = w; B.Us[4][g] = p;
}

// Stable-step radius max over nodes of r_c + r_v, from the RK base U0 = det Q (valid between
// steps); per-workgroup maxima into part[], then one workgroup reduces them into part[nparts].
//   r_c = sum_d |u . grad xi_d| + c |grad xi_d|                         (convection)
//   r_v = VISC_RADIUS * mu / (rho Re) * (sum_d |grad xi_d|)^2            (viscous blocks only)
// r_v bounds the diffusion operator of the viscous terms: the 4th-order central first derivative
// (unit computational spacing) has a symbol of modulus at most 1.372, applied twice along any pair
// of directions through the metric rows, times the largest diffusivity factor max(4/3, 1.4 / Pr)
// of the stresses and the heat flux; mu is Sutherland's law at the node temperature (the viscous
// kernel's), Re, Pr, T_inf and the kernel Mach number those of the block's parameters.
constexpr int RED = 256;
constexpr double VISC_RADIUS = 1.372 * 1.372;

__device__ __forceinline__ double block_max(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = RED / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    return sh[0];
}

__global__ void __launch_bounds__(RED) k_cfl_part(BlockDev B, double *part)
{
    __shared__ double sh[RED];
    double r = 0.0;
    const uint32_t no = (uint32_t)B.ni * B.nj * B.nk;
    for (uint32_t o = blockIdx.x * RED + threadIdx.x; o < no; o += gridDim.x * RED) {
        const uint32_t i = o % (uint32_t)B.ni, j = (o / (uint32_t)B.ni) % (uint32_t)B.nj, k = o / ((uint32_t)B.ni * B.nj);
        const uint32_t g = (i + B.off[0]) + (uint32_t)B.pi * ((j + B.off[1]) + (uint32_t)B.pj * (k + B.off[2]));
        const double idet = 1.0 / B.det[g];
        const double rho = idet * B.U0[0][g], ir = 1.0 / rho;
        const double u = (idet * B.U0[1][g]) * ir, v = (idet * B.U0[2][g]) * ir, w = (idet * B.U0[3][g]) * ir;
        const double p = 0.4 * ((idet * B.U0[4][g]) - 0.5 * rho * (u * u + v * v + w * w));
        const double c = sqrt(1.4 * p / rho);
        double s = 0.0, sm = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double mx = B.m[3 * d][g], my = B.m[3 * d + 1][g], mz = B.m[3 * d + 2][g];
            const double mm = sqrt(mx * mx + my * my + mz * mz);
            s += fabs(u * mx + v * my + w * mz) + c * mm;
            sm += mm;
        }
        if (B.viscous) {
            const cfdx::dev::ViscParams &vp = B.vp;
            const double T = vp.tfac * (p / (0.4 * rho));          // 1.4 M^2 0.4 e
            const double mu = (T * sqrt(T)) * (vp.suth1 / (T + vp.suth));
            const double diff = fmax(4.0 / 3.0, 1.4 / vp.pr);
            s += VISC_RADIUS * diff * (mu * vp.inv_reyl / rho) * (sm * sm);
        }
        r = fmax(r, s);
    }
    const double m = block_max(r, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void __launch_bounds__(RED) k_cfl_final(double *part, int nparts)
{
    __shared__ double sh[RED];
    double r = 0.0;
    for (int q = threadIdx.x; q < nparts; q += RED) r = fmax(r, part[q]);
    const double m = block_max(r, sh);
    if (threadIdx.x == 0) part[nparts] = m;
}

}  // namespace

void init_geometry(const BlockDev &B, const double L[3], double amp, const double ph[3], hipStream_t st)
{
    MapArgs a{{L[0], L[1], L[2]}, amp, {ph[0], ph[1], ph[2]}};
    const unsigned nb = (unsigned)((B.n + 255) / 256);
    hipLaunchKernelGGL(k_geom, dim3(nb), dim3(256), 0, st, B, a);
    hipLaunchKernelGGL(k_half, dim3(nb), dim3(256), 0, st, B);
}

void init_state(const BlockDev &B, const double L[3], double amp, const double ph[3], int kind, double mach, double beta,
                double xc, double yc, hipStream_t st)
{
    InitArgs a{{{L[0], L[1], L[2]}, amp, {ph[0], ph[1], ph[2]}}, kind, mach, beta, xc, yc};
    const unsigned nb = (unsigned)((B.n + 255) / 256);
    hipLaunchKernelGGL(k_init, dim3(nb), dim3(256), 0, st, B, a);
}

int cfl_parts(const BlockDev &B) { return (int)std::min<uint32_t>(1024u, ((uint32_t)B.ni * B.nj * B.nk + RED - 1) / RED); }

void cfl_radius(const BlockDev &B, double *part, hipStream_t st)
{
    const int np = cfl_parts(B);
    hipLaunchKernelGGL(k_cfl_part, dim3(np), dim3(RED), 0, st, B, part);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(RED), 0, st, part, np);
}

}  // namespace blk
}  // namespace cfdx
