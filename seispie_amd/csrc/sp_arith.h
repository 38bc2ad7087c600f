// seispie_hip -- the finite-difference arithmetic shared by the step
// kernels, in exact and fast form.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

// float32 operations the reference performs in float32 (float32 * float32
// stays float32 under Numba's typing); contraction off so they are never
// folded into a float64 multiply-add.
__device__ __forceinline__ float mul_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a - b;
}

// D- at row r from values at r-2, r-1, r, r+1 (solver.py:22-27, 38-43)
__device__ __forceinline__ double dminus(float m2, float m1, float c0, float p1, double c1, double c2)
{
    return c1 * (double)sub_rn(c0, m1) - c2 * (double)sub_rn(p1, m2);
}
// D+ at row q from values at q-1, q, q+1, q+2 (solver.py:30-35, 46-51)
__device__ __forceinline__ double dplus(float m1, float c0, float p1, float p2, double c1, double c2)
{
    return c1 * (double)sub_rn(p1, c0) - c2 * (double)sub_rn(p2, m1);
}

// 1/x in float64 to within an ulp: hardware estimate + two Newton steps
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// ---- exact mode: the reference's float64 operations, one rounding each, in
// its order (no contraction), so results equal the oracle bit for bit.
// a / c correctly rounded from rc = RN(1/c): q = RN(a rc), the remainder
// a - q c is exact by FMA, and RN(q + r rc) is the correctly rounded quotient
// (Markstein's correction step; also checked against IEEE division on
// 2e8 random operands of these ranges).
__device__ __forceinline__ double rn_div(double a, double c, double rc)
{
#pragma clang fp contract(off)
    const double q = a * rc;
    const double r = __builtin_fma(-q, c, a);
    return __builtin_fma(r, rc, q);
}
// RN(1/x) for float32 x: three Newton steps from the hardware estimate.  Two
// already give the correctly rounded reciprocal of every float32 mantissa
// from any start within 2^-20 (exhaustive check); the third keeps it so from
// starts as far off as 2^-14.
__device__ __forceinline__ double rcp_rn(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
// RN(num x / den) as fma(x, hi, x lo) with num/den split into hi + lo (the
// host checks the identity for every float32 mantissa of x)
__device__ __forceinline__ double kmul(double x, double hi, double lo)
{
    return __builtin_fma(x, hi, x * lo);
}
// D- and D+ as the reference writes them: 9 (a) / (8h) - (b) / (24h)
__device__ __forceinline__ double dminus_x(float m2, float m1, float c0, float p1, double k9h, double k9l,
                                           double k1h, double k1l)
{
#pragma clang fp contract(off)
    return kmul((double)sub_rn(c0, m1), k9h, k9l) - kmul((double)sub_rn(p1, m2), k1h, k1l);
}
__device__ __forceinline__ double dplus_x(float m1, float c0, float p1, float p2, double k9h, double k9l,
                                          double k1h, double k1l)
{
#pragma clang fp contract(off)
    return kmul((double)sub_rn(p1, c0), k9h, k9l) - kmul((double)sub_rn(p2, m1), k1h, k1l);
}
// add_vy / add_vxz: v = bound (v + dt ds / rho), then u = u + v dt
__device__ __forceinline__ float vel_x(float vo, float ds, double dt, double rho, double irho, double b)
{
#pragma clang fp contract(off)
    return (float)(b * ((double)vo + rn_div(dt * (double)ds, rho, irho)));
}
__device__ __forceinline__ float disp_x(float u, float v, double dt)
{
#pragma clang fp contract(off)
    return (float)((double)u + (double)v * dt);
}
// s + f * g, one rounding per operation (add_sy and the spin couple stresses)
__device__ __forceinline__ float sadd_x(float s, double f, double g)
{
#pragma clang fp contract(off)
    return (float)((double)s + f * g);
}
// add_sxz: s + dt (l2m d1 + float32(lam d2))
__device__ __forceinline__ float snorm_x(float s, double dt, double l2m, float d1, float lam_d2)
{
#pragma clang fp contract(off)
    return (float)((double)s + dt * (l2m * (double)d1 + (double)lam_d2));
}
// add_sy_c: syy_c + 2 dt nu (-vy_c - 0.5 float32(dvzdx - dvxdz))
__device__ __forceinline__ float srot_x(float s, double dt2, float nu, float vyc, float curl)
{
#pragma clang fp contract(off)
    return (float)((double)s + dt2 * (double)nu * (-(double)vyc - 0.5 * (double)curl));
}

// the updates in either mode (fast: contraction allowed)
template <bool EX>
__device__ __forceinline__ float upd_sadd(float s, double f, double g)
{
    if constexpr (EX) return sadd_x(s, f, g);
    else return (float)((double)s + f * g);
}
template <bool EX>
__device__ __forceinline__ float upd_snorm(float s, double dt, double l2m, float d1, float lam_d2)
{
    if constexpr (EX) return snorm_x(s, dt, l2m, d1, lam_d2);
    else return (float)((double)s + dt * (l2m * (double)d1 + (double)lam_d2));
}
template <bool EX>
__device__ __forceinline__ float upd_srot(float s, double dt2, float nu, float vyc, float curl)
{
    if constexpr (EX) return srot_x(s, dt2, nu, vyc, curl);
    else return (float)((double)s + dt2 * (double)nu * (-(double)vyc - 0.5 * (double)curl));
}
template <bool EX>
__device__ __forceinline__ float upd_disp(float u, float v, double dt)
{
    if constexpr (EX) return disp_x(u, v, dt);
    else return (float)((double)u + (double)v * dt);
}
// v where the lane's condition holds, else 0.  The argument is evaluated
// unconditionally, so a per-lane condition becomes a select instead of an
// exec-masked branch around the stencil (scalar instructions and a break in
// the scheduling region per use).
__device__ __forceinline__ double lsel(bool ok, double v) { return ok ? v : 0.0; }
__device__ __forceinline__ float lsel(bool ok, float v) { return ok ? v : 0.f; }
// A wave-uniform condition the compiler has to treat as per-lane: its uses
// become selects of unconditionally computed values (one v_cndmask each)
// instead of scalar branches, each of which ends a scheduling region.
__device__ __forceinline__ bool lane_bool(bool b)
{
    int v = b;
    asm volatile("" : "+v"(v));
    return v != 0;
}

// derivative coefficients of a shot in both forms
struct DCo {
    double c1, c2, k9h, k9l, k1h, k1l;
};
template <bool EX>
__device__ __forceinline__ double dmn(float m2, float m1, float c0, float p1, const DCo &k)
{
    if constexpr (EX) return dminus_x(m2, m1, c0, p1, k.k9h, k.k9l, k.k1h, k.k1l);
    else return dminus(m2, m1, c0, p1, k.c1, k.c2);
}
template <bool EX>
__device__ __forceinline__ double dpl(float m1, float c0, float p1, float p2, const DCo &k)
{
    if constexpr (EX) return dplus_x(m1, c0, p1, p2, k.k9h, k.k9l, k.k1h, k.k1l);
    else return dplus(m1, c0, p1, p2, k.c1, k.c2);
}
