// Device physics shared by the drop-in kernels and the structured sweeps.
// T is the arithmetic type of fluxes / reconstruction / viscous terms (double for
// the plain form, float for the mixed form). gamma = 1.4 as the reference's
// literals 1.4 / 0.4 / 0.2 (Weno.cl:2292-2295, U_Q.cl:73).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace cfdx {
namespace dev {

template <typename T> __device__ __forceinline__ T sqrt_(T x);
template <> __device__ __forceinline__ double sqrt_<double>(double x) { return __builtin_sqrt(x); }
template <> __device__ __forceinline__ float sqrt_<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T> __device__ __forceinline__ T fabs_(T x) { return x < T(0) ? -x : x; }

// Reciprocal and quotient: exact IEEE division in double (the plain form's parity bar), the
// hardware reciprocal / square root (about 1 ulp) in float for the mixed form.
template <typename T> __device__ __forceinline__ T inv_(T x);
template <> __device__ __forceinline__ double inv_<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float inv_<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename T> __device__ __forceinline__ T div_(T a, T b);
template <> __device__ __forceinline__ double div_<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float div_<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Reciprocal, square root and reciprocal square root of a strictly positive normal operand (a
// density, a squared norm, a sum of positive terms): the hardware estimate (relative error below
// 6e-8 on gfx950) refined in double by one Newton-Raphson step (reciprocal: relative error below
// 2.2e-15, 0.7 ulp on average) or one Goldschmidt step and one residual correction (square root:
// within 0.5 ulp, correctly rounded for 99.98 % of operands; its reciprocal within 1.5 ulp); the
// hardware instructions in float (within 1 ulp). Measured on 2^20 log-uniform operands in
// [1e-300, 1e300]. Not for 0, inf, NaN or operands whose result is subnormal (no range scaling).
template <typename T> __device__ __forceinline__ T pinv_(T x);
template <> __device__ __forceinline__ double pinv_<double>(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
template <> __device__ __forceinline__ float pinv_<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// s = sqrt(x) and h = 1 / (2 sqrt(x)) by one Goldschmidt step on the rsq estimate, then one
// residual correction of s
__device__ __forceinline__ double gold_sqrt(double x, double &h)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
template <typename T> __device__ __forceinline__ T psqrt_(T x);
template <> __device__ __forceinline__ double psqrt_<double>(double x) { double h; return gold_sqrt(x, h); }
template <> __device__ __forceinline__ float psqrt_<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
// 1 / sqrt(x)
template <typename T> __device__ __forceinline__ T prsq_(T x);
template <> __device__ __forceinline__ double prsq_<double>(double x)
{
    double h;
    const double s = gold_sqrt(x, h);
    const double r = 2.0 * h;
    return __builtin_fma(r, __builtin_fma(-s, r, 1.0), r);   // one Newton step on 1 / s
}
template <> __device__ __forceinline__ float prsq_<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// Conserved state Q = (rho, rho u, rho v, rho w, p/0.4 + rho |u|^2 / 2) of primitives (U_Q.cl:118-122).
// No fused multiply-adds, so every kernel that builds the RK base det * Q from the same primitives
// (the last RK stage, set_flow) gets the same bits: a state saved and restored continues exactly.
template <typename T>
__device__ __forceinline__ void prim_to_q(T rho, T u, T v, T w, T p, T (&q)[5])
{
#pragma clang fp contract(off)
    const T ke = T(0.5) * rho * (u * u + v * v + w * w);
    q[0] = rho;
    q[1] = rho * u;
    q[2] = rho * v;
    q[3] = rho * w;
    q[4] = p * T(2.5) + ke;   // p / (gamma - 1)
}

// Total enthalpy H = (E + p) / rho of primitives (U_Q.cl:73-74).
template <typename T>
__device__ __forceinline__ T enthalpy(T rho, T u, T v, T w, T p)
{
    return (p * T(3.5) + T(0.5) * rho * (u * u + v * v + w * w)) * pinv_(rho);   // (p/0.4 + ke + p) / rho
}

// Roe-average inputs (rho, u, v, w, H) of a node from its conserved state Q: the face kernels hold
// the Q of the stencil points already, so the two centre nodes need no primitive loads.
template <typename T>
__device__ __forceinline__ void q_roe_inputs(const T (&q)[5], T &rho, T &u, T &v, T &w, T &h)
{
    const T r = pinv_(q[0]);
    rho = q[0];
    u = q[1] * r;
    v = q[2] * r;
    w = q[3] * r;
    h = (T(1.4) * q[4] - T(0.2) * (q[1] * u + q[2] * v + q[3] * w)) * r;   // (E + p) / rho, p = 0.4 (E - rho |u|^2 / 2)
}

// Roe average of the face between states a and b (U_Q.cl:60-109): (u, v, w, h, rho).
template <typename T>
__device__ __forceinline__ void roe_average(T ra, T ua, T va, T wa, T ha, T rb, T ub, T vb, T wb, T hb, T (&r)[5])
{
    const T sa = psqrt_(ra), sb = psqrt_(rb);
    const T is = pinv_(sa + sb);
    r[0] = (sb * ub + sa * ua) * is;
    r[1] = (sb * vb + sa * va) * is;
    r[2] = (sb * wb + sa * wa) * is;
    r[3] = (sb * hb + sa * ha) * is;
    r[4] = sb * sa;
}

// Contravariant, det-scaled inviscid node flux along metric row m (Viscous.cl:560-611):
// D * (mx F1 + my E1 + mz G1) with F1, E1, G1 the Cartesian Euler fluxes; Ep = E + p.
template <typename T>
__device__ __forceinline__ void node_flux_e(T rho, T u, T v, T w, T p, T Ep, T mx, T my, T mz, T D, T (&f)[5])
{
    const T V = mx * u + my * v + mz * w;
    const T rV = rho * V;
    f[0] = D * rV;
    f[1] = D * (rV * u + mx * p);
    f[2] = D * (rV * v + my * p);
    f[3] = D * (rV * w + mz * p);
    f[4] = D * (Ep * V);
}
template <typename T>
__device__ __forceinline__ void node_flux(T rho, T u, T v, T w, T p, T mx, T my, T mz, T D, T (&f)[5])
{
    const T V = mx * u + my * v + mz * w;
    const T Ep = p * T(3.5) + T(0.5) * rho * (u * u + v * v + w * w);   // E + p
    const T rV = rho * V;
    f[0] = D * rV;
    f[1] = D * (rV * u + mx * p);
    f[2] = D * (rV * v + my * p);
    f[3] = D * (rV * w + mz * p);
    f[4] = D * (Ep * V);
}

// WCNS-E5 interpolation of one characteristic field to the face between W[2] and W[3]
// (W[0..5] at offsets -2..+3 from the left node), Weno.cl:1056-1279:
//   candidates (15/8, -5/4, 3/8), (3/8, 3/4, -1/8), (-1/8, 3/4, 3/8),
//   Jiang-Shu indicators, w_k = d_k / (eps + IS_k)^2, value = sum w_k q_k / sum w_k.
// pos = right-biased (d = 1/16, 5/8, 5/16), neg = left-biased (d = 5/16, 5/8, 1/16).
template <typename T>
__device__ __forceinline__ void wcns_pair(const T (&W)[6], T eps, T &pos, T &neg)
{
    // the candidates, the second differences and the indicators' first-derivative terms are the
    // same linear forms of W, written in the first differences d_m = W[m+1] - W[m]
    const T d0 = W[1] - W[0], d1 = W[2] - W[1], d2 = W[3] - W[2], d3 = W[4] - W[3], d4 = W[5] - W[4];
    const T h_r = W[3] + T(3.0 / 8.0) * d4 - T(7.0 / 8.0) * d3;    // pos q0 = (15 W3 - 10 W4 + 3 W5) / 8
    const T h_c = W[3] - T(3.0 / 8.0) * d2 - T(1.0 / 8.0) * d3;    // pos q1 = neg q0 = (3 W2 + 6 W3 - W4) / 8
    const T h_l = W[2] + T(1.0 / 8.0) * d1 + T(3.0 / 8.0) * d2;    // pos q2 = neg q1 = (-W1 + 6 W2 + 3 W3) / 8
    const T h_ll = W[2] + T(7.0 / 8.0) * d1 - T(3.0 / 8.0) * d0;   // neg q2 = (3 W0 - 10 W1 + 15 W2) / 8
    // the indicators are formed four times larger, 4 IS = 13/3 s^2 + t^2, from the second
    // differences s_k = d_{k+1} - d_k (each 13/3 s_k^2 shared by two indicators) and the
    // first-derivative terms t; with eps also taken four times larger every (eps + IS)^2 carries the
    // same factor 16, which cancels in the weights
    const T c13 = T(13.0 / 3.0), eps4 = T(4) * eps;
    const T s0 = d1 - d0, s1 = d2 - d1, s2 = d3 - d2, s3 = d4 - d3;
    // eps + 13/3 s_k^2, each shared by two indicators
    const T S0 = c13 * s0 * s0 + eps4, S1 = c13 * s1 * s1 + eps4, S2 = c13 * s2 * s2 + eps4, S3 = c13 * s3 * s3 + eps4;
    const T t0 = d4 - T(3) * d3;   // 3 W3 - 4 W4 + W5
    const T t1 = d2 + d3;          // W4 - W2
    const T t2 = T(3) * d2 - d1;   // W1 - 4 W2 + 3 W3
    const T u0 = d3 - T(3) * d2;   // 3 W2 - 4 W3 + W4
    const T u1 = d1 + d2;          // W3 - W1
    const T u2 = T(3) * d1 - d0;   // W0 - 4 W1 + 3 W2
    const T ISp0 = S3 + t0 * t0, ISp1 = S2 + t1 * t1, ISp2 = S1 + t2 * t2;   // eps + IS, four times larger
    const T ISm0 = S2 + u0 * u0, ISm1 = S1 + u1 * u1, ISm2 = S0 + u2 * u2;
    // w_k proportional to d_k * prod_{m != k} (eps + IS_m)^2, so each side needs one division; the
    // linear weights are taken 16 times larger (pos 1, 10, 5; neg 5, 10, 1): the factor cancels in
    // the quotients
    T np_, dp_, nn_, dn_;
    {
        T a0 = ISp0, a1 = ISp1, a2 = ISp2;
        a0 *= a0; a1 *= a1; a2 *= a2;
        const T e = a1 * a2, A2 = T(10) * a2, A1 = T(5) * a1;
        np_ = e * h_r + a0 * (A2 * h_c + A1 * h_l);
        dp_ = e + a0 * (A2 + A1);
    }
    {
        T b0 = ISm0, b1 = ISm1, b2 = ISm2;
        b0 *= b0; b1 *= b1; b2 *= b2;
        const T e = b0 * b1, B1 = T(5) * b1, B0 = T(10) * b0;
        nn_ = e * h_ll + b2 * (B1 * h_c + B0 * h_l);
        dn_ = e + b2 * (B1 + B0);
    }
    if constexpr (!std::is_same<T, float>::value) {
        // one reciprocal serves both sides
        const T r = pinv_(dp_ * dn_);
        pos = np_ * (dn_ * r);
        neg = nn_ * (dp_ * r);
    } else {
        // float: the weights of smooth data are ~eps^4 ~ 1e-28, within the float range, but the
        // product of the two sides' sums (~eps^8) is not: one hardware reciprocal per side
        pos = np_ * inv_(dp_);
        neg = nn_ * inv_(dn_);
    }
}

// Characteristic reconstruction of the two interface states of one face
// (Weno.cl:336-455 eigenvectors and projection, :1056-2187 WCNS, :2188-2287 back projection).
//   q[m][c]: conserved Q at offsets -2..+3 from the left node
//   roe = (u, v, w, h, rho) of the face; (kx, ky, kz) = the face metric row over its length
//   qp / qn: right-biased / left-biased conserved states (not det-scaled)
template <typename T>
__device__ __forceinline__ void characteristic_states(const T (&q)[6][5], const T (&roe)[5], T kx, T ky, T kz, T eps,
                                                      T (&qp)[5], T (&qn)[5])
{
    const T u = roe[0], v = roe[1], w = roe[2], h = roe[3], rho = roe[4];
    const T q2 = u * u + v * v + w * w;
    const T a = psqrt_(T(0.4) * (h - T(0.5) * q2));
    const T theta = kx * u + ky * v + kz * w;
    const T phi = T(0.2) * q2;
    const T irho = pinv_(rho);
    const T ia = pinv_(a);
    const T iaa = ia * ia;
    const T s2 = T(0.70710678118654752440);           // 1/sqrt(2)
    const T alpha = rho * ia * s2;                      // rho / (sqrt2 a)
    const T beta = irho * ia * s2;                      // 1 / (sqrt2 rho a)
    const T g4aa = T(0.4) * iaa;
    const T one_m = T(1) - phi * iaa;
    const T cv = kz * v - ky * w, cw = kx * w - kz * u, cu = ky * u - kx * v;  // u x k
    const T phi_g = phi * T(2.5);
    const T hh = (phi + a * a) * T(2.5);

    // Projection W = L q of the 6 stencil states, written through L's structure: rows 0-2 are
    // k_c E + (1/rho) (q_mom x k - (u x k) q0)_c with E = (1 - phi/a^2) q0 + (0.4/a^2) m and
    // m = u.q_mom - q4; rows 3 and 4 are beta (phi q0 - 0.4 m) -/+ beta a (theta q0 - k.q_mom).
    const T ikx = irho * kx, iky = irho * ky, ikz = irho * kz;
    const T icu = irho * cu, icv = irho * cv, icw = irho * cw;
    const T bphi = beta * phi, b04 = beta * T(0.4), ba = beta * a;
    T W[5][6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const T q0 = q[m][0], qx = q[m][1], qy = q[m][2], qz = q[m][3], q4 = q[m][4];
        const T mm = u * qx - q4 + v * qy + w * qz;
        const T nn = kx * qx + ky * qy + kz * qz;
        const T E = one_m * q0 + g4aa * mm;
        W[0][m] = kx * E + (ikz * qy - iky * qz - icv * q0);
        W[1][m] = ky * E + (ikx * qz - ikz * qx - icw * q0);
        W[2][m] = kz * E + (iky * qx - ikx * qy - icu * q0);
        const T bA = bphi * q0 - b04 * mm;
        const T Bv = theta * q0 - nn;
        W[3][m] = bA - ba * Bv;
        W[4][m] = bA + ba * Bv;
    }
    T xp[5], xn[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) wcns_pair(W[c], eps, xp[c], xn[c]);
    // Back projection Q = R x through R's structure: with x' = (x0, x1, x2), X = k.x',
    // P = alpha (x3 + x4) and M = alpha a (x3 - x4): Q0 = X + P, Q_mom = u Q0 + k M + rho (k x x'),
    // Q4 = (phi/0.4) X + H P + theta M + rho (u x k).x'
    const T rkx = rho * kx, rky = rho * ky, rkz = rho * kz;
    const T rcu = rho * cu, rcv = rho * cv, rcw = rho * cw;
    const T aa = alpha * a;
    auto back = [&](const T (&x)[5], T (&Q)[5]) {
        const T X = kx * x[0] + ky * x[1] + kz * x[2];
        const T P = alpha * (x[3] + x[4]);
        const T M = aa * (x[3] - x[4]);
        Q[0] = X + P;
        Q[1] = u * Q[0] + kx * M + rky * x[2] - rkz * x[1];
        Q[2] = v * Q[0] + ky * M + rkz * x[0] - rkx * x[2];
        Q[3] = w * Q[0] + kz * M + rkx * x[1] - rky * x[0];
        Q[4] = phi_g * X + hh * P + theta * M + rcv * x[0] + rcw * x[1] + rcu * x[2];
    };
    back(xp, qp);
    back(xn, qn);
}

// n . (F, E, G)(Q) and lambda_5 = n.u + c of one interface state along the unit face normal n
// (Weno.cl:2291-2507 fluxes, :2511-2581 eigenvalues, :2585-2655 contravariant form: the face's
// |k| det_f scales the LLF sum once).
template <typename T>
__device__ __forceinline__ void face_state_flux(const T (&Q)[5], T kx, T ky, T kz, T (&c)[5], T &lam)
{
    const T r = pinv_(Q[0]);
    const T ke = (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * r;
    const T P = T(0.4) * (Q[4] - T(0.5) * ke);
    const T Un = (kx * Q[1] + ky * Q[2] + kz * Q[3]) * r;
    c[0] = Q[0] * Un;
    c[1] = Q[1] * Un + kx * P;
    c[2] = Q[2] * Un + ky * P;
    c[3] = Q[3] * Un + kz * P;
    c[4] = (T(1.4) * Q[4] - T(0.2) * ke) * Un;
    lam = Un + psqrt_(T(1.4) * P * r);
}

// LLF face flux with alpha = max(|lambda_5|) of the two interface states (Weno.cl:2859-2947):
// (kx, ky, kz) the unit face normal, dk = det_f |k| of the face metric row k.
template <typename T>
__device__ __forceinline__ void face_llf(const T (&q)[6][5], const T (&roe)[5], T kx, T ky, T kz, T dk, T eps,
                                         T (&llf)[5])
{
    T qp[5], qn[5];
    characteristic_states(q, roe, kx, ky, kz, eps, qp, qn);
    T cp[5], cn[5], lp, ln;
    face_state_flux(qp, kx, ky, kz, cp, lp);
    face_state_flux(qn, kx, ky, kz, cn, ln);
    const T al = fabs_(lp) >= fabs_(ln) ? fabs_(lp) : fabs_(ln);
    const T hd = T(0.5) * dk;
#pragma unroll
    for (int c = 0; c < 5; ++c) llf[c] = hd * (cp[c] + cn[c] - al * (qp[c] - qn[c]));
}

// Explicit midpoint correction - delta2 F / 24 + 7 delta4 F / 5760 of one component from the node
// fluxes at offsets -2..+3 (Weno.cl:2951-3058).
template <typename T>
__device__ __forceinline__ T midpoint_correction(T f0, T f1, T f2, T f3, T f4, T f5)
{
    // d2 = (-5 f0 + 39 f1 - 34 f2 - 34 f3 + 39 f4 - 5 f5) / 48,
    // d4 = (f0 - 3 f1 + 2 f2 + 2 f3 - 3 f4 + f5) / 2; -d2 / 24 + 7 d4 / 5760 collects to
    // (57 (f0 + f5) - 411 (f1 + f4) + 354 (f2 + f3)) / 11520
    return T(57.0 / 11520.0) * (f0 + f5) - T(411.0 / 11520.0) * (f1 + f4) + T(354.0 / 11520.0) * (f2 + f3);
}

// Face flux F-hat = LLF + midpoint correction from node fluxes nf at offsets -2..+3; (k0, k1, k2)
// the face metric row as stored by the reference, detf the face det.
template <typename T, typename TO>
__device__ __forceinline__ void face_flux(const T (&q)[6][5], const T (&roe)[5], T k0, T k1, T k2, T detf, T eps,
                                          const T (&nf)[6][5], TO (&out)[5])
{
    const T kn = sqrt_(k0 * k0 + k1 * k1 + k2 * k2), ik = inv_(kn);
    T llf[5];
    face_llf(q, roe, k0 * ik, k1 * ik, k2 * ik, detf * kn, eps, llf);
#pragma unroll
    for (int c = 0; c < 5; ++c)
        out[c] = TO(llf[c] + midpoint_correction(nf[0][c], nf[1][c], nf[2][c], nf[3][c], nf[4][c], nf[5][c]));
}

// Viscous + SGS node fluxes (Viscous.cl:439-631, its mixed-time-scale model or the Smagorinsky
// option) of one node.
//   dU[c][d] = dU_c / d(zeta, eta, xi)_d of the det-scaled state U (Viscous.cl:285-435)
//   U[5]     = U at the node; u, v, w = the velocity at the node (Viscous.cl:146-154 reads it from
//              FLOW; the block kernel passes U_1..3 / U_0 of the same stage vector away from the
//              sides of a block)
//   m[9]     = node metrics zeta_x..xi_z; D = det
//   KS[3]    = MTS scale-similarity terms KS_x, KS_y, KS_z (Viscous.cl:493-495)
// Writes the contravariant det-scaled viscous fluxes fvis[0..2] = Fv, Ev, Gv and returns mu.
struct ViscParams {
    double mach, reyl, pr, free_t, c_mts, c_t;
    // derived on the host (uniform per launch)
    double inv_reyl;     // 1 / Re
    double kheat;        // 1.4 / (Pr Re)
    double suth;         // 110.4 / T_inf
    double suth1;        // 1 + 110.4 / T_inf
    double tfac;         // 1.4 M^2 0.4
    double kQ;           // -1 / (1.4 0.4 M^2)
    double inv_ct;       // 1 / C_T
    double cs2;          // Smagorinsky: C_s^2
    int sgs;             // 0: mixed time scale (the reference); 1: Smagorinsky
    double kheat_t;      // Smagorinsky heat flux: 1.4 / Pr
};

inline ViscParams make_visc_params(double mach, double reyl, double pr, double free_t, double c_mts, double c_t,
                                   int sgs = 0, double c_s = 0.0)
{
    ViscParams v{mach, reyl, pr, free_t, c_mts, c_t, 0, 0, 0, 0, 0, 0, 0, c_s * c_s, sgs, 1.4 / pr};
    v.inv_reyl = 1.0 / reyl;
    v.kheat = 1.4 / (pr * reyl);
    v.suth = 110.4 / free_t;
    v.suth1 = 1.0 + v.suth;
    v.tfac = 1.4 * mach * mach * 0.4;
    v.kQ = -(1.0 / (1.4 * 0.4 * mach * mach));
    v.inv_ct = 1.0 / c_t;
    return v;
}

// x^(-1/3) of a positive normal operand (a det): the float estimate 2^(-log2(x) / 3) of the
// hardware log2 and exp2, refined by one Newton step y <- y + y (1 - x y^3) / 3 in float and two
// more in double (within 1 ulp on operands in [1e-36, 1e36]; float form within 1 ulp too).
template <typename T> __device__ __forceinline__ T prcbrt_(T x);
template <> __device__ __forceinline__ float prcbrt_<float>(float x)
{
    const float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (-1.0f / 3.0f));
    return __builtin_fmaf(y * (1.0f / 3.0f), __builtin_fmaf(-x * y, y * y, 1.0f), y);
}
template <> __device__ __forceinline__ double prcbrt_<double>(double x)
{
    double y = (double)prcbrt_<float>((float)x);
    y = __builtin_fma(y * (1.0 / 3.0), __builtin_fma(-x * y, y * y, 1.0), y);
    return __builtin_fma(y * (1.0 / 3.0), __builtin_fma(-x * y, y * y, 1.0), y);
}

template <typename T>
__device__ __forceinline__ T viscous_node(const T (&dU)[5][3], const T (&U)[5], T u, T v, T w, const T (&m)[9], T D,
                                          const T (&KS)[3], const ViscParams &vp, T (&fvis)[3][5])
{
    const T C_MTS = T(vp.c_mts);
    const T iU0 = pinv_(U[0]);
    const T iU02 = iU0 * iU0;
    T ug[3], vg[3], wg[3], tg[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ug[d] = iU0 * dU[1][d] - (U[1] * iU02) * dU[0][d];
        vg[d] = iU0 * dU[2][d] - (U[2] * iU02) * dU[0][d];
        wg[d] = iU0 * dU[3][d] - (U[3] * iU02) * dU[0][d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
        tg[d] = iU0 * dU[4][d] - (U[4] * iU02) * dU[0][d] - (U[1] * iU0) * ug[d] - (U[2] * iU0) * vg[d] - (U[3] * iU0) * wg[d];
    // Cartesian gradients: X_x = X_zeta zeta_x + X_eta eta_x + X_xi xi_x, etc.
    const T ux = ug[0] * m[0] + ug[1] * m[3] + ug[2] * m[6];
    const T uy = ug[0] * m[1] + ug[1] * m[4] + ug[2] * m[7];
    const T uz = ug[0] * m[2] + ug[1] * m[5] + ug[2] * m[8];
    const T vx = vg[0] * m[0] + vg[1] * m[3] + vg[2] * m[6];
    const T vy = vg[0] * m[1] + vg[1] * m[4] + vg[2] * m[7];
    const T vz = vg[0] * m[2] + vg[1] * m[5] + vg[2] * m[8];
    const T wx = wg[0] * m[0] + wg[1] * m[3] + wg[2] * m[6];
    const T wy = wg[0] * m[1] + wg[1] * m[4] + wg[2] * m[7];
    const T wz = wg[0] * m[2] + wg[1] * m[5] + wg[2] * m[8];
    const T ex = tg[0] * m[0] + tg[1] * m[3] + tg[2] * m[6];
    const T ey = tg[0] * m[1] + tg[1] * m[4] + tg[2] * m[7];
    const T ez = tg[0] * m[2] + tg[1] * m[5] + tg[2] * m[8];

    // Viscous.cl:455-458: T = gamma M^2 (gamma-1) e, Sutherland mu with S = 110.4
    const T d_temp = T(vp.tfac) * (U[4] * iU0 - T(0.5) * ((U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * iU02));
    const T st = psqrt_(d_temp);
    const T mu = (st * st * st) * (T(vp.suth1) * pinv_(d_temp + T(vp.suth)));
    // Viscous.cl:461-476 heat flux and Stokes stresses
    const T kq = -(mu * T(vp.kheat));
    const T qx = kq * ex, qy = kq * ey, qz = kq * ez;
    const T mr = mu * T(vp.inv_reyl);
    const T txx = mr * (T(4.0 / 3.0) * ux - T(2.0 / 3.0) * vy - T(2.0 / 3.0) * wz);
    const T tyy = mr * (T(4.0 / 3.0) * vy - T(2.0 / 3.0) * ux - T(2.0 / 3.0) * wz);
    const T tzz = mr * (T(4.0 / 3.0) * wz - T(2.0 / 3.0) * ux - T(2.0 / 3.0) * vy);
    const T txy = mr * (uy + vx);
    const T txz = mr * (uz + wx);
    const T tyz = mr * (vz + wy);
    // Viscous.cl:490-558 mixed-time-scale eddy viscosity, one per direction
    const T sxy = T(0.5) * (uy + vx), syz = T(0.5) * (vz + wy), sxz = T(0.5) * (wx + uz);
    const T sxx = ux, syy = vy, szz = wz;
    const T mod = sqrt_(T(2) * ((sxy * sxy + syz * syz + sxz * sxz) + T(2.0 / 3.0) * (sxx * sxx + syy * syy + szz * szz)));
    const T DEL = prcbrt_(D);
    T eddy[3];
    if (vp.sgs == 1) {
        // Smagorinsky (no reference counterpart): nu_t = (C_s Delta)^2 |S| with Delta = det^(1/3) and
        // |S| = sqrt(2 S_ij S_ij), the same in every direction. The SGS stress below is
        // -2 U0 eddy S_ij with U0 = det rho, so eddy = nu_t / det = C_s^2 |S| det^(-1/3) = C_s^2 |S| DEL.
        const T s2 = T(2) * (sxx * sxx + syy * syy + szz * szz) + T(4) * (sxy * sxy + syz * syz + sxz * sxz);
        const T e = T(vp.cs2) * DEL * sqrt_(s2);
        eddy[0] = e; eddy[1] = e; eddy[2] = e;
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T sk = sqrt_(KS[d]);
            if (mod != T(0)) {
                // C_MTS / (1 + R / C_T) * DEL * sk with R = DEL |S| / sk, written with one division
                // (sk = 0 gives R = inf and a zero eddy viscosity either way)
                eddy[d] = (C_MTS * DEL * sk * sk) * pinv_(sk + DEL * mod * T(vp.inv_ct));
            } else {
                eddy[d] = C_MTS * DEL * sk;
            }
        }
    }
    // SGS stresses scale with U0 = det * rho (Viscous.cl:529-535)
    const T m2U = T(-2) * U[0];
    const T Sxy = m2U * eddy[0] * sxy, Syz = m2U * eddy[1] * syz, Sxz = m2U * eddy[2] * sxz;
    const T Sxx = m2U * eddy[0] * sxx, Syy = m2U * eddy[1] * syy, Szz = m2U * eddy[2] * szz;
    T Hx, Hy, Hz;
    if (vp.sgs == 1) {
        // Smagorinsky: the molecular forms with rho nu_t (= U0 eddy) for mu / Re: heat flux
        // -1.4 rho nu_t / Pr grad e, and the work of the SGS stresses
        const T kS = -(T(vp.kheat_t) * (U[0] * eddy[0]));
        Hx = kS * ex + (u * Sxx + v * Sxy + w * Sxz);
        Hy = kS * ey + (u * Sxy + v * Syy + w * Syz);
        Hz = kS * ez + (u * Sxz + v * Syz + w * Szz);
    } else {
        const T kQ = T(vp.kQ);
        const T Qx = kQ * eddy[0] * ex, Qy = kQ * eddy[1] * ey, Qz = kQ * eddy[2] * ez;
        const T DX = (Sxx + Sxy + Sxz) * u, DY = (Sxy + Syy + Syz) * v, DZ = (Sxz + Syz + Szz) * w;
        // H = Q + DEL and D = 0.4 Q (Viscous.cl:549-557)
        Hx = Qx + DX + T(0.4) * Qx; Hy = Qy + DY + T(0.4) * Qy; Hz = Qz + DZ + T(0.4) * Qz;
    }
    // Viscous.cl:576-595 Cartesian viscous fluxes
    const T tF[5] = {T(0), -txx + Sxx, -txy + Sxy, -txz + Sxz, (-u * txx - v * txy - w * txz) + qx + Hx};
    const T tE[5] = {T(0), -txy + Sxy, -tyy + Syy, -tyz + Syz, (-u * txy - v * tyy - w * tyz) + qy + Hy};
    const T tG[5] = {T(0), -txz + Sxz, -tyz + Syz, -tzz + Szz, (-u * txz - v * tyz - w * tzz) + qz + Hz};
    // Viscous.cl:615-631 contravariant and det-scaled
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 5; ++c) fvis[d][c] = D * (m[3 * d] * tF[c] + m[3 * d + 1] * tE[c] + m[3 * d + 2] * tG[c]);
    return mu;
}

// One-direction derivative of Viscous.cl:285-435 (weno_form = false) or Weno.cl:2732-2856
// (weno_form = true) selected by the loc codes; f[0..6] at offsets -3..+3. Branches are
// taken in source order, a later match overwriting an earlier one; with no match the
// reference leaves the value undefined and this returns 0.
template <typename T>
__device__ __forceinline__ T loc_derivative(const T (&f)[7], int LI, int LM2, int LM1, int LP1, int LP2, bool weno_form)
{
    T d = T(0);
    if (LI > 0 && LI <= 52 && LP1 == 100) d = T(1.0 / 6.0) * (T(11) * f[3] - T(18) * f[2] + T(9) * f[1] - T(2) * f[0]);
    if (LI > 0 && LI <= 52 && LM1 == 100) d = T(1.0 / 6.0) * (T(-11) * f[3] + T(18) * f[4] - T(9) * f[5] + T(2) * f[6]);
    if (LI == 0 && LP1 > 0 && LP1 <= 52 && LP2 == 100) d = T(1.0 / 6.0) * (T(2) * f[4] + T(3) * f[3] - T(6) * f[2] + f[1]);
    if (LI == 0 && LM1 > 0 && (weno_form ? LM1 <= 52 : true) && LM2 == 100)
        d = T(1.0 / 6.0) * (T(-2) * f[2] - T(3) * f[3] + T(6) * f[4] - f[5]);
    if (LM2 != 100 && LM1 != 100 && LP2 != 100 && LP1 != 100 && (LI == 0 || (!weno_form && LI > 0)))
        d = T(1.0 / 12.0) * (T(8) * (f[4] - f[2]) - (f[5] - f[1]));
    return d;
}

}  // namespace dev
}  // namespace cfdx
