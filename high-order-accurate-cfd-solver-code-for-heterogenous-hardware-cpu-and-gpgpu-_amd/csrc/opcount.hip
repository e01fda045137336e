// Static operation count of the scheme's arithmetic: the templated physics of the face and node
// kernels (physics.hpp) instantiated with a counting number type and run once, by one thread, on
// representative data (the central branches; no data-dependent loop). Each +, -, * of a value
// counts one operation (a fused multiply-add counts two), as do /, reciprocal, square root and
// cube root; operations between compile-time constants do not count. The double-precision
// arithmetic written out in the node kernels themselves (stencil differences, the RK update,
// primitive recovery) is counted by hand below, from block_kernels.hip.
#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "physics.hpp"

namespace cfdx {
namespace opc {

struct Ops {
    unsigned long long add, mul, div, sqrt, rcp, cbrt;
};

// one counting thread; its tallies live in device memory
__device__ Ops g_ops;

struct Cnt {
    double v;
    bool var;   // a value computed from the inputs (not a compile-time constant)
    __device__ Cnt() : v(0.0), var(false) {}
    __device__ Cnt(double x) : v(x), var(false) {}
    __device__ Cnt(double x, bool t) : v(x), var(t) {}
};

__device__ inline Cnt tally(unsigned long long &slot, double v, bool var)
{
    if (var) ++slot;
    return Cnt(v, var);
}
__device__ inline Cnt operator+(Cnt a, Cnt b) { return tally(g_ops.add, a.v + b.v, a.var || b.var); }
__device__ inline Cnt operator-(Cnt a, Cnt b) { return tally(g_ops.add, a.v - b.v, a.var || b.var); }
__device__ inline Cnt operator*(Cnt a, Cnt b) { return tally(g_ops.mul, a.v * b.v, a.var || b.var); }
__device__ inline Cnt operator/(Cnt a, Cnt b) { return tally(g_ops.div, a.v / b.v, a.var || b.var); }
__device__ inline Cnt operator-(Cnt a) { return Cnt(-a.v, a.var); }
__device__ inline Cnt &operator+=(Cnt &a, Cnt b) { return a = a + b; }
__device__ inline Cnt &operator-=(Cnt &a, Cnt b) { return a = a - b; }
__device__ inline Cnt &operator*=(Cnt &a, Cnt b) { return a = a * b; }
__device__ inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
__device__ inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
__device__ inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
__device__ inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }
__device__ inline bool operator!=(Cnt a, Cnt b) { return a.v != b.v; }

}  // namespace opc

namespace dev {
using opc::Cnt;
using opc::g_ops;
template <> __device__ Cnt sqrt_<Cnt>(Cnt x) { return opc::tally(g_ops.sqrt, __builtin_sqrt(x.v), x.var); }
template <> __device__ Cnt psqrt_<Cnt>(Cnt x) { return opc::tally(g_ops.sqrt, __builtin_sqrt(x.v), x.var); }
template <> __device__ Cnt prsq_<Cnt>(Cnt x) { return opc::tally(g_ops.sqrt, 1.0 / __builtin_sqrt(x.v), x.var); }
template <> __device__ Cnt inv_<Cnt>(Cnt x) { return opc::tally(g_ops.rcp, 1.0 / x.v, x.var); }
template <> __device__ Cnt pinv_<Cnt>(Cnt x) { return opc::tally(g_ops.rcp, 1.0 / x.v, x.var); }
template <> __device__ Cnt div_<Cnt>(Cnt a, Cnt b) { return opc::tally(g_ops.div, a.v / b.v, a.var || b.var); }
template <> __device__ Cnt prcbrt_<Cnt>(Cnt x) { return opc::tally(g_ops.cbrt, 1.0 / cbrt(x.v), x.var); }
}  // namespace dev

namespace opc {
using namespace cfdx::dev;

__device__ inline Cnt in(double x) { return Cnt(x, true); }

__device__ inline void take(Ops &dst)
{
    dst = g_ops;
    g_ops = Ops{0, 0, 0, 0, 0, 0};
}

// out[0]: one face of a face sweep (Roe state from the two centre nodes, characteristic WCNS
// pair, face fluxes, LLF, 6th-order midpoint correction of the 5 components);
// out[1]: conserved state of a node; out[2]: node flux along one metric row;
// out[3]: viscous + MTS node fluxes of a node
__global__ void k_opcount(Ops *out, ViscParams vp)
{
    g_ops = Ops{0, 0, 0, 0, 0, 0};
    {
        Cnt q[6][5], nf[6][5];
        for (int m = 0; m < 6; ++m) {
            const double s = 1.0 + 0.01 * m;
            q[m][0] = in(1.0 * s); q[m][1] = in(0.9 * s); q[m][2] = in(0.1 * s); q[m][3] = in(0.05 * s); q[m][4] = in(3.2 * s);
            for (int c = 0; c < 5; ++c) nf[m][c] = in(0.3 * (c + 1) * s);
        }
        // the FP64 sweep takes the Roe inputs of the two centre nodes from their conserved states
        Cnt ra, ua, va, wa, ha, rb, ub, vb, wb, hb;
        q_roe_inputs(q[2], ra, ua, va, wa, ha);
        q_roe_inputs(q[3], rb, ub, vb, wb, hb);
        Cnt roe[5];
        roe_average(ra, ua, va, wa, ha, rb, ub, vb, wb, hb, roe);
        Cnt llf[5];
        // the unit face normal and det_f |k| as the sweeps read them
        face_llf(q, roe, in(0.99), in(0.11), in(0.055), in(1.2), Cnt(1e-7), llf);
        Cnt fh[5];
        for (int c = 0; c < 5; ++c)
            fh[c] = llf[c] + midpoint_correction(nf[0][c], nf[1][c], nf[2][c], nf[3][c], nf[4][c], nf[5][c]);
        take(out[0]);
    }
    {
        Cnt q[5];
        prim_to_q(in(1.0), in(0.9), in(0.1), in(0.05), in(1.1), q);
        take(out[1]);
        // the sweeps form E + p from the conserved energy just computed
        Cnt f[5];
        node_flux_e(in(1.0), in(0.9), in(0.1), in(0.05), in(1.1), q[4] + in(1.1), in(0.9), in(0.1), in(0.05), in(1.2), f);
        take(out[2]);
    }
    {
        Cnt dU[5][3], U[5], m[9], KS[3], fv[3][5];
        for (int c = 0; c < 5; ++c) {
            U[c] = in(1.0 + 0.1 * c);
            for (int d = 0; d < 3; ++d) dU[c][d] = in(0.01 * (c + 1) * (d + 2));
        }
        for (int c = 0; c < 9; ++c) m[c] = in(c % 4 == 0 ? 1.0 : 0.05);
        for (int d = 0; d < 3; ++d) KS[d] = in(1e-4 * (d + 1));
        viscous_node(dU, U, in(0.9), in(0.1), in(0.05), m, in(1.2), KS, vp, fv);
        take(out[3]);
    }
}

}  // namespace opc
}  // namespace cfdx

// Operations written out in double in the node kernels (block_kernels.hip), counted by hand:
//   visc_node: per direction and component the central difference (8 (f+1 - f-1) - (f+2 - f-2))
//     / 12 = 5 operations (75 for 5 components x 3 directions); per direction the MTS
//     scale-similarity term a = U/U0 - (weighted mean)/(weighted mean), a^2 = 14 operations
//     (42); 1 / U0 = 1; the node velocity U_m / U_0 away from the sides = 3 (total 121).
//   update_node: face-flux differences 2 x 15 = 30; viscous divergence 4 components x 3
//     directions x (5 + 1) = 72; RK stage 5 x {2, 6, 6} (modes 0, 1, 2; mean 4.67) = 23.3 per
//     stage; primitive recovery (1/det, rho, 1/rho, u v w, e, p) = 20; at the last stage the RK
//     base det * Q = the conserved-state count + 5.
extern "C" int cfdx_op_count(cfdx_ctx *ctx, const cfdx_params *params, int64_t out[24])
{
    CFDX_TRY(cfdx::enter(ctx));
    if (!out) return cfdx::fail(CFDX_E_INVALID, "op_count: null output");
    cfdx_params d;
    if (!params) { cfdx_params_default(&d); params = &d; }
    const cfdx::dev::ViscParams vp =
        cfdx::dev::make_visc_params(params->mach, params->reyl, params->pr, params->free_t, params->c_mts, params->c_t,
                                    params->sgs, params->c_s);
    cfdx::opc::Ops *dev = nullptr;
    CFDX_HIP(hipMalloc((void **)&dev, 4 * sizeof(cfdx::opc::Ops)), "op_count hipMalloc");
    hipLaunchKernelGGL(cfdx::opc::k_opcount, dim3(1), dim3(1), 0, ctx->stream, dev, vp);
    int rc = cfdx::check_launch("k_opcount");
    cfdx::opc::Ops host[4];
    if (rc == CFDX_OK) {
        hipError_t e = hipMemcpyAsync(host, dev, sizeof host, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = cfdx::hip_fail(e, "op_count copy");
    }
    (void)hipFree(dev);
    if (rc != CFDX_OK) return rc;
    for (int g = 0; g < 4; ++g) {
        const cfdx::opc::Ops &o = host[g];
        const int64_t v[6] = {(int64_t)o.add, (int64_t)o.mul, (int64_t)o.div, (int64_t)o.sqrt, (int64_t)o.rcp, (int64_t)o.cbrt};
        for (int c = 0; c < 6; ++c) out[6 * g + c] = v[c];
    }
    return CFDX_OK;
}

namespace cfdx {
namespace opc {

// y[t] = the device helper fn at x[t]: 0..3 the double forms of the positive-operand reciprocal,
// square root, reciprocal square root and inverse cube root, 4..7 their float forms (operand
// rounded to float, result widened)
__global__ void __launch_bounds__(256) k_math_eval(int fn, const double *__restrict__ x, double *__restrict__ y,
                                                   int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double v = x[t];
    const float f = (float)v;
    double r = 0.0;
    switch (fn) {
    case 0: r = dev::pinv_<double>(v); break;
    case 1: r = dev::psqrt_<double>(v); break;
    case 2: r = dev::prsq_<double>(v); break;
    case 3: r = dev::prcbrt_<double>(v); break;
    case 4: r = (double)dev::pinv_<float>(f); break;
    case 5: r = (double)dev::psqrt_<float>(f); break;
    case 6: r = (double)dev::prsq_<float>(f); break;
    default: r = (double)dev::prcbrt_<float>(f); break;
    }
    y[t] = r;
}

}  // namespace opc
}  // namespace cfdx

extern "C" int cfdx_math_eval(cfdx_ctx *ctx, int fn, const double *x, double *y, int64_t n)
{
    CFDX_TRY(cfdx::enter(ctx));
    if (fn < 0 || fn > 7) return cfdx::fail(CFDX_E_INVALID, "math_eval: fn must be 0..7");
    if (n < 0 || n > (int64_t(1) << 24)) return cfdx::fail(CFDX_E_INVALID, "math_eval: n must be 0..2^24");
    if (n == 0) return CFDX_OK;
    if (!x || !y) return cfdx::fail(CFDX_E_INVALID, "math_eval: null array");
    double *dev = nullptr;
    const size_t bytes = (size_t)n * sizeof(double);
    CFDX_HIP(hipMalloc((void **)&dev, 2 * bytes), "math_eval hipMalloc");
    hipError_t e = hipMemcpyAsync(dev, x, bytes, hipMemcpyHostToDevice, ctx->stream);
    int rc = e == hipSuccess ? CFDX_OK : cfdx::hip_fail(e, "math_eval copy in");
    if (rc == CFDX_OK) {
        hipLaunchKernelGGL(cfdx::opc::k_math_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, fn, dev,
                           dev + n, n);
        rc = cfdx::check_launch("k_math_eval");
    }
    if (rc == CFDX_OK) {
        e = hipMemcpyAsync(y, dev + n, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = cfdx::hip_fail(e, "math_eval copy out");
    }
    (void)hipFree(dev);
    return rc;
}
