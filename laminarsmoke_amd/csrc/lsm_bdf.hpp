// lsm_bdf.hpp - the OpenSMOKE++ Gear/BDF driver as one device template.
//
// The step, order and Newton logic of
//   math/native-ode-solvers/MultiValueSolver.hpp:424-1328 (Solve, DriverStep,
//   NewOrderNewH, MultiValueStep, Interpolation, CalculateInitialStepSize),
//   math/native-ode-solvers/MethodGear.hpp:80-536 (ResetMethod, FindCorrection,
//   WhatToDoInCaseOfConvergenceFailure, ConvergenceRate, CheckConstraints)
// is written once against an execution context C.  The context owns the
// Nordsieck/work vectors and the linear algebra and decides how a cell maps
// to lanes:
//   * TeamCtx  - one cell per wavefront, one vector element per lane, the
//                Newton matrix factorised in registers (lsm_team.hpp);
//   * PlainCtx - one cell per thread, vectors in HBM (lsm_plain.hpp).
// All scalars below are uniform across the lanes that work on one cell.
//
// Context interface:
//   int ne;                         number of equations (ns + 1)
//   template<class F> void each(F)  f(e) for every element slot e
//   double& x(int vid, int e)       element e of vector vid (vid compile-time)
//   double ymin(e), ymax(e)         chemistry_DI.H:82-83 bounds (+-inf on padding)
//   double reduce(double partial)   sum over the elements of a cell (uniform result)
//   int count(bool)                 number of elements for which the flag holds
//   void rhs(int src, int dst)      dst = f(src); updates c.qr
//   int jacobian(double h)          Jacobian at VB (FD with f at F, weights OOE, or
//                                   analytic); returns the RHS evaluations it used
//   void factor(double hr0)         G = I - hr0 J, LU
//   void solve(int vid)             in place G^-1
//   void jac_times(int src, int dst)
#pragma once

// Phase timers of the BDF bookkeeping (builds with LSM_PHASE_TIMERS only): the
// context's add_time(slot, cycles) accumulates them.
#ifdef LSM_PHASE_TIMERS
#define LSM_BDF_TIC(v) const unsigned long long v = clock64()
#define LSM_BDF_TOC(c, i, v) (c).add_time(i, clock64() - (v))
#else
#define LSM_BDF_TIC(v)
#define LSM_BDF_TOC(c, i, v)
#endif

namespace lsm {

enum VecId {
    Z0 = 0,   // Z0..Z7 Nordsieck array z_[0..7]
    V0 = 8,   // V0..V7 working array v_[0..7]
    VB_ = 16, VA_ = 17, DB_ = 18, F_ = 19, B_ = 20, OOE_ = 21, Y_ = 22, Y0_ = 23, AUX_ = 24,
    NVEC = 25
};

constexpr int MAX_ORDER = 5;
constexpr int MIN_ORDER = 1;
// OpenSMOKEFunctions.hpp:39-40,60-85 and OpenSMOKEStdInclude.h:53-54
constexpr double EPS_D = 2.220446049250313e-16;
constexpr double EPS_F = 1.1920928955078125e-07;
constexpr double TINY_F = 1.1754943508222875e-38;
constexpr double BIG_D = 1.7976931348623154e+308;

enum { H_DEC = 0, H_CONST = 1, H_INC = 2 };
enum { O_DEC = 0, O_CONST = 1, O_INC = 2 };
enum { J_CHANGE = 0, J_MODIFIED = 1, J_OK = 2 };
enum { M_TO_FACTOR = 0, M_FACTORED = 1 };
enum { ST_INIT = 1, ST_CONT = 2, ST_MAXSTEPS = -1, ST_STRICT = -2, ST_ERRFAIL = -4,
       ST_CONVFAIL = -5, ST_SMALLH = -6, ST_ILLEGAL = -9 };

struct BdfConst {
    double r[MAX_ORDER + 1][MAX_ORDER];   // MethodGear.hpp:155-178 (computed on the host)
    double abs_tol, rel_tol;
    double t0;
    unsigned max_steps;
};

struct BdfCounters {
    unsigned steps, fcalls, fcalls_jac, jacobians, factorizations, solves;
    unsigned conv_failures, err_failures, max_order, rhs_total, clamps, newton;
};

struct BdfState {
    double t, t0, tOut, tInMeshPoint, tStabilize;
    double h, hNextStep, hNordsieck, hPreviousStep, hScale, hScaleMax, first_step;
    double alfa2_4, alfa2_5;     // alfa2_[4], alfa2_[5]; the other entries stay 0.8
    __device__ __forceinline__ double alfa2(int i) const { return i == 4 ? alfa2_4 : (i == 5 ? alfa2_5 : 0.8); }
    int p, orderInNextStep, maximum_order, status;
    int odeHStatus, odeOrderStatus, factorizationStatus, jacobianStatus, convOK;
    unsigned iterConvergence, iterConvergenceRate, iterConvergenceFailure, iterErrorFailure, iterOrder;
    unsigned maxIterationsJacobian, stepOfLastJacobian;
    unsigned convFailOrderMax, iterConstSteps, iterMinimumOrder;
    int fixedOrderMin;
};

__device__ __forceinline__ double ep_fact(int j) {   // Ep_[j] = j! (MethodGear.hpp:150-152)
    double f = 1.;
    for (int i = 2; i <= j; i++) f *= (double)i;
    return f;
}

// WRMS norm of vector vid (ErrorControl, OpenSMOKEUtilities.hpp:689-701)
template <class C>
__device__ __forceinline__ double wrms(C &c, int vid) {
    double acc = 0.;
    c.each([&](int e) { const double t = c.x(vid, e) * c.x(OOE_, e); acc += t * t; });
    return sqrt(c.reduce(acc) / (double)c.ne);
}

// WRMS of v_[q] for a runtime q (registers need compile-time vector ids)
template <class C>
__device__ __forceinline__ double wrms_v(C &c, int q) {
    double r = 0.;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (k == q) r = wrms(c, V0 + k);
    return r;
}

template <class C>
__device__ __forceinline__ void check_constraints(C &c, int vid, BdfCounters &n) {
    int cnt = 0;
    c.each([&](int e) {
        double &y = c.x(vid, e);
        const bool lo = y < c.ymin(e), hi = y > c.ymax(e);
        if (lo) y = c.ymin(e);
        if (hi) y = c.ymax(e);
        cnt += c.count(lo || hi);
    });
    n.clamps += cnt;
}

// CalculatesOneOverEpsilon, MultiValueSolver.hpp:620-646 (scalar tolerances)
template <class C>
__device__ __forceinline__ void one_over_epsilon(C &c, const BdfConst &K, int vid) {
    c.each([&](int e) { c.x(OOE_, e) = 1. / (K.abs_tol + K.rel_tol * fabs(c.x(vid, e))); });
}

// MethodGear::FindCorrection, MethodGear.hpp:196-341
template <class C>
__device__ __forceinline__ unsigned find_correction(C &c, const BdfConst &K, BdfState &s, BdfCounters &n) {
    c.each([&](int e) { c.x(VB_, e) = c.x(V0, e); });
    check_constraints(c, VB_, n);
    c.rhs(VB_, F_);
    n.fcalls++;
    n.rhs_total++;
    if (n.steps >= s.stepOfLastJacobian + s.maxIterationsJacobian) s.jacobianStatus = J_CHANGE;
    if (s.jacobianStatus == J_CHANGE) {
        s.stepOfLastJacobian = n.steps;
        s.jacobianStatus = J_MODIFIED;
        const int nj = c.jacobian(s.h);   // FD: ne RHS evaluations; analytic: none
        n.fcalls_jac += nj;
        n.rhs_total += nj;
        n.jacobians++;
        s.factorizationStatus = M_TO_FACTOR;
    }
    const double r0 = K.r[0][s.p - 1];
    const double hr0 = s.h * r0;
    if (s.factorizationStatus == M_FACTORED && s.iterOrder == 0) s.factorizationStatus = M_TO_FACTOR;
    if (s.factorizationStatus == M_TO_FACTOR) {
        n.factorizations++;
        c.factor(hr0);
        s.factorizationStatus = M_FACTORED;
    }
    c.each([&](int e) { c.x(DB_, e) = c.x(F_, e) * s.h - c.x(V0 + 1, e); });
    c.solve(DB_);
    n.solves++;
    c.each([&](int e) { c.x(B_, e) = c.x(DB_, e); });
    double old = 0.;
    for (unsigned k = 1; k <= 4; k++) {   // maxConvergenceIterations_ = 4 (MethodGear.hpp:45)
        const double cc = r0 * wrms(c, DB_);
        if (cc < 1.) { s.convOK = 1; n.newton += k; return k; }
        if (k > 1 && cc > 2. * old) { s.convOK = 0; n.newton += k; return k; }
        old = cc;
        int cl = 0;
        c.each([&](int e) {
            double va = c.x(B_, e) * r0 + c.x(V0, e);
            const bool lo = va < c.ymin(e);
            if (lo) { va = c.ymin(e); c.x(B_, e) = (c.ymin(e) - c.x(Z0, e)) / r0; }
            const bool hi = va > c.ymax(e);
            if (hi) { va = c.ymax(e); c.x(B_, e) = (c.ymax(e) - c.x(Z0, e)) / r0; }
            c.x(VA_, e) = va;
            cl += c.count(lo || hi);
        });
        n.clamps += cl;
        c.rhs(VA_, F_);
        n.fcalls++;
        n.rhs_total++;
        c.each([&](int e) { c.x(DB_, e) = c.x(F_, e) * s.h - c.x(V0 + 1, e) - c.x(B_, e); });
        c.solve(DB_);
        n.solves++;
        c.each([&](int e) { c.x(B_, e) += c.x(DB_, e); });
    }
    s.convOK = 0;
    n.newton += 4;
    return 4;
}

// MultiValueSolver::MultiValueStep, MultiValueSolver.hpp:1213-1328
template <class C>
__device__ __forceinline__ void multi_value_step(C &c, const BdfConst &K, BdfState &s, BdfCounters &n) {
    const int p = s.p;
    // Updates under the order p are written as selects, not branches: with
    // branches the compiler moves part of the Nordsieck array to scratch memory.
    c.each([&](int e) {
#pragma unroll
        for (int i = 0; i <= MAX_ORDER; i++) {
            const double z = c.x(Z0 + i, e), v = c.x(V0 + i, e);
            c.x(V0 + i, e) = (i <= p) ? z : v;
        }
#pragma unroll
        for (int i = 0; i < MAX_ORDER; i++)
#pragma unroll
            for (int j = MAX_ORDER - 1; j >= i; j--) {
                const double a = c.x(V0 + j, e), b = c.x(V0 + j + 1, e);
                c.x(V0 + j, e) = (i < p && j <= p - 1) ? a + b : a;
            }
    });
    const double dne = (double)c.ne;
    int cl = 0;
    c.each([&](int e) {
        const double v0 = c.x(V0, e);
        const double lo = c.ymin(e), hi = c.ymax(e);
        const bool clo = v0 < lo && fabs(v0 - lo) * c.x(OOE_, e) < dne;
        const bool chi = v0 > hi && fabs(v0 - hi) * c.x(OOE_, e) < dne;
        const bool cc = clo || chi;
        const double bnd = clo ? lo : hi;
        const double v1 = c.x(V0 + 1, e), z0 = c.x(Z0, e);
        c.x(V0, e) = cc ? bnd : v0;
        c.x(V0 + 1, e) = cc ? bnd - z0 : v1;
#pragma unroll
        for (int i = 2; i <= MAX_ORDER; i++) {
            const double vi = c.x(V0 + i, e);
            c.x(V0 + i, e) = (cc && i <= p) ? 0. : vi;
        }
        cl += c.count(clo || chi);
    });
    n.clamps += cl;
    s.iterConvergence = find_correction(c, K, s, n);
    if (!s.convOK) return;
    const double r0 = K.r[0][p - 1];
    double rr[MAX_ORDER + 1];
#pragma unroll
    for (int i = 0; i <= MAX_ORDER; i++) rr[i] = (i >= 2 && i <= p) ? K.r[i][p - 1] : 0.;
    const double inv_p1 = 1. / ((double)p + 1.), inv_p2 = 1. / ((double)p + 2.);
    const bool do_p2 = s.iterOrder >= (unsigned)p;
    cl = 0;
    c.each([&](int e) {
        double va = c.x(B_, e) * r0;
        const double lo = c.ymin(e), hi = c.ymax(e);
        const bool clo = c.x(V0, e) + va < lo;
        if (clo) {
            c.x(B_, e) = lo - c.x(Z0, e) - c.x(V0 + 1, e);
            double v = lo - va;
            c.x(V0, e) = v < lo ? lo : v;
        }
        const bool chi = c.x(V0, e) + va > hi;
        if (chi) {
            c.x(B_, e) = hi - c.x(Z0, e) - c.x(V0 + 1, e);
            double v = hi - va;
            c.x(V0, e) = v > hi ? hi : v;
        }
        cl += c.count(clo || chi);
        const double b = c.x(B_, e);
        c.x(V0, e) += va;
        c.x(V0 + 1, e) += b;
#pragma unroll
        for (int i = 2; i <= MAX_ORDER; i++) {
            const double vi = c.x(V0 + i, e);
            c.x(V0 + i, e) = (i <= p) ? vi + b * rr[i] : vi;
        }
        // error estimates v_[p+1], v_[p+2] (eq. 29.169)
        double vp = 0., zp = 0., zp1 = 0.;
#pragma unroll
        for (int i = 1; i <= MAX_ORDER; i++) {
            const double a = c.x(V0 + i, e), z = c.x(Z0 + i, e), z1 = c.x(Z0 + i + 1, e);
            vp = (i == p) ? a : vp;
            zp = (i == p) ? z : zp;
            zp1 = (i == p) ? z1 : zp1;
        }
        const double vp1 = (vp - zp) / ((double)p + 1.);
        (void)inv_p1;
        (void)inv_p2;
#pragma unroll
        for (int i = 2; i <= MAX_ORDER + 1; i++) {
            const double vi = c.x(V0 + i, e);
            c.x(V0 + i, e) = (i == p + 1) ? vp1 : vi;
        }
        if (do_p2) {
            const double vp2 = (vp1 - zp1) / ((double)p + 2.);
#pragma unroll
            for (int i = 3; i <= MAX_ORDER + 2; i++) {
                const double vi = c.x(V0 + i, e);
                c.x(V0 + i, e) = (i == p + 2) ? vp2 : vi;
            }
        }
    });
    n.clamps += cl;
}

// MethodGear::WhatToDoInCaseOfConvergenceFailure, MethodGear.hpp:344-423
template <class C>
__device__ __forceinline__ void convergence_failure(C &c, BdfState &s, BdfCounters &n) {
    s.hScale = 0.25;
    if (n.steps < 10) s.hScale = 0.1;
    s.hScale = fmax(s.hScale, 0. / fabs(s.h));   // min_step_size_ = 0 (ResetMethod)
    s.odeHStatus = H_DEC;
    if (s.p > MIN_ORDER) { s.p -= 1; s.odeOrderStatus = O_DEC; }
    s.alfa2_4 -= 0.01;
    s.alfa2_5 -= 0.02;
    if (s.alfa2_4 < 0.) {
        s.alfa2_4 = 0.; s.maximum_order = 3;
        if (s.p > 3) { s.p = 3; s.odeOrderStatus = O_DEC; }
    }
    if (s.alfa2_5 < 0.) {
        s.alfa2_5 = 0.; s.maximum_order = 4;
        if (s.p > 4) { s.p = 4; s.odeOrderStatus = O_DEC; }
    }
    if (n.steps > 5) {
        s.convFailOrderMax++;
        if (s.convFailOrderMax % 5 == 0 && s.p != MIN_ORDER) { s.p = MIN_ORDER; s.odeOrderStatus = O_DEC; }
        if (s.convFailOrderMax % 10 == 0 && s.tStabilize != s.tInMeshPoint) {
            c.rhs(Z0, VA_);
            n.fcalls++;
            n.rhs_total++;
            c.each([&](int e) { c.x(Z0 + 1, e) = c.x(VA_, e) * s.h; });
            s.tStabilize = s.tInMeshPoint;
        }
    }
    if (n.steps > s.stepOfLastJacobian + 50 / 10) s.jacobianStatus = J_CHANGE;
    if (n.steps != s.stepOfLastJacobian && s.iterConvergenceFailure == 50) s.jacobianStatus = J_CHANGE;
    s.hNextStep *= s.hScale;
}

// MethodGear::ConvergenceRate, MethodGear.hpp:426-516
template <class C>
__device__ __forceinline__ void convergence_rate(C &c, BdfState &s, BdfCounters &n) {
    if (s.iterErrorFailure > 0 || s.iterConvergenceFailure > 0) s.iterConvergenceRate = 0;
    if (s.iterConvergence <= 1) return;
    c.each([&](int e) { c.x(VB_, e) = c.x(V0, e) - c.x(Z0, e); });
    c.jac_times(VB_, VA_);
    double pd = 0., pf = 0.;
    c.each([&](int e) {
        const double va = c.x(V0 + 1, e) - (c.x(VA_, e) * s.h + c.x(Z0 + 1, e));
        c.x(VA_, e) = va;
        pd += fabs(va);
        pf += fabs(c.x(V0 + 1, e));
    });
    const double normd = c.reduce(pd), normf = c.reduce(pf);
    if (normd > 3. * normf + .1 && s.iterConvergence > 2) {
        if (n.steps > s.stepOfLastJacobian + 50 / 10) s.jacobianStatus = J_CHANGE;
        s.hScale = 0.5; s.odeHStatus = H_DEC; s.hNextStep = s.h * s.hScale;
        if (s.p > 1) { s.orderInNextStep = s.p - 1; s.odeOrderStatus = O_DEC; }
        s.iterConvergenceRate = 0; s.iterConvergenceFailure = 1;
        return;
    }
    if (normd > normf + 0.1) {
        s.iterConvergenceRate++;
        if (s.iterConvergenceRate >= 5) {
            s.hScale = 0.8; s.odeHStatus = H_DEC; s.hNextStep = s.h * s.hScale;
            s.iterConvergenceRate = 0; s.iterConvergenceFailure = 1;
        }
        if (n.steps > s.stepOfLastJacobian + 50 / 10 && s.iterConvergence > 2) {
            s.jacobianStatus = J_CHANGE;
            s.iterConvergenceRate = 0; s.iterConvergenceFailure = 1;
        }
    }
}

// MultiValueSolver::NewOrderNewH, MultiValueSolver.hpp:1043-1210.  The CPU-time
// heuristics never fire in the reference (MultiValueSolver::cpuTimeToFactorize_
// is reset to 0 and never updated, :133), so scl = 1.2.
template <class C>
__device__ __forceinline__ void new_order_new_h(C &c, BdfState &s, const BdfCounters &n) {
    const int p = s.p;
    double hm = 0., hc = 0., hp = 0., alfa1 = 0., alfa3 = 0.;
    if (p > MIN_ORDER && s.odeOrderStatus != O_INC) {
        hm = ep_fact(p - 1) * wrms_v(c, p) + EPS_F;
        hm = 1. / hm;
        alfa1 = s.alfa2(p - 1) - 0.05;
        hm = alfa1 * pow(hm, 1. / (double)p);
    }
    hc = ep_fact(p) * wrms_v(c, p + 1) + EPS_F;
    hc = 1. / hc;
    hc = s.alfa2(p) * pow(hc, 1. / (double)(p + 1));
    if (p < s.maximum_order && s.odeOrderStatus != O_DEC) {
        hp = ep_fact(p + 1) * wrms_v(c, p + 2) + EPS_F;
        hp = 1. / hp;
        alfa3 = s.alfa2(p + 1) - 0.25;
        hp = alfa3 * pow(hp, 1. / (double)(p + 2));
    }
    if (hm > hc && hm > hp) {
        s.hScale = (.75 / alfa1) * hm; s.orderInNextStep = p - 1; s.odeOrderStatus = O_DEC;
    } else if (hc > hm && hc > hp) {
        s.hScale = (.75 / s.alfa2(p)) * hc; s.orderInNextStep = p; s.odeOrderStatus = O_CONST;
    } else {
        s.hScale = (.75 / alfa3) * hp; s.orderInNextStep = p + 1; s.odeOrderStatus = O_INC;
    }
    if (s.fixedOrderMin) {
        s.iterMinimumOrder++;
        if (s.iterMinimumOrder < 2000) {
            s.hScale = (.75 / s.alfa2(p)) * hc; s.orderInNextStep = p; s.odeOrderStatus = O_CONST;
        } else s.fixedOrderMin = 0;
    }
    if (s.hScale < 1.2) {
        s.iterConstSteps++;
        s.hScale = 1.; s.odeHStatus = H_CONST; s.orderInNextStep = p; s.odeOrderStatus = O_CONST;
        s.hNextStep = s.h;
    } else {
        s.iterConstSteps = 0;
        s.hScale = fmin(s.hScale, s.hScaleMax);
        s.odeHStatus = H_INC;
        s.hNextStep = s.h * s.hScale;
        if (s.hNextStep > BIG_D) s.hNextStep = BIG_D;
    }
    if (s.iterConstSteps > 500 ||
        (n.steps % 1000 == 0 && n.steps != 0 && (fabs(s.h) < 1.e-7 || fabs(s.h) < 1.e-5 * s.t))) {
        if (p > MIN_ORDER) {
            s.fixedOrderMin = 1; s.iterMinimumOrder = 0;
            s.orderInNextStep = MIN_ORDER; s.odeOrderStatus = O_DEC;
            s.hScale = 0.5; s.hNextStep = s.h * s.hScale; s.odeHStatus = H_DEC;
        }
    }
}

// MultiValueSolver::DriverStep, MultiValueSolver.hpp:754-1040
template <class C>
__device__ __forceinline__ void driver_step(C &c, const BdfConst &K, BdfState &s, BdfCounters &n) {
    s.iterConvergenceFailure = 0; s.iterErrorFailure = 0; s.convOK = 1;
    if (s.odeOrderStatus != O_CONST || s.odeHStatus != H_CONST) s.iterOrder = 0;
    s.p = s.orderInNextStep;
    s.odeOrderStatus = O_CONST;
    s.hScaleMax = 2.;
    if (n.steps < 10) s.hScaleMax = 1000.;
    else if (n.steps > 10 && n.steps < 20) s.hScaleMax = 10.;
    for (;;) {
        if (s.odeHStatus != H_CONST) {
            s.hScale = s.hNextStep / s.hNordsieck;
            const double hs = s.hScale;
            const int p = s.p;
            c.each([&](int e) {
                double hr = 1.;
#pragma unroll
                for (int i = 1; i <= MAX_ORDER; i++) {
                    if (i <= p) { hr *= hs; c.x(Z0 + i, e) *= hr; }
                }
            });
            check_constraints(c, Z0, n);
            s.iterOrder = 0;
            s.odeHStatus = H_CONST;
        }
        s.h = s.hNextStep;
        s.hNordsieck = s.hNextStep;
        if (fabs(s.h) < EPS_F * 0.) { s.status = ST_SMALLH; return; }   // min_step_size_ = 0
        s.t = s.tInMeshPoint + s.h;
        if (s.t == s.tInMeshPoint) { s.status = ST_SMALLH; return; }
        LSM_BDF_TIC(t_mvs);
        multi_value_step(c, K, s, n);
        LSM_BDF_TOC(c, 10, t_mvs);
        LSM_BDF_TIC(t_tail);
        if (!s.convOK) {
            s.iterConvergenceFailure++;
            n.conv_failures++;
            if (s.iterConvergenceFailure >= 50) { s.status = ST_CONVFAIL; return; }
            if (fabs(s.h - 0.) <= EPS_D * 0.) { s.status = ST_SMALLH; return; }
            convergence_failure(c, s, n);
        } else {
            double hsc = ep_fact(s.p) * wrms_v(c, s.p + 1);
            if (hsc > 1.) {
                s.iterErrorFailure++;
                n.err_failures++;
                if (s.iterErrorFailure >= 50) { s.status = ST_ERRFAIL; return; }
                if (fabs(s.h - 0.) <= EPS_D * 0.) { s.status = ST_SMALLH; return; }
                s.hScaleMax = 1.;
                hsc = 0.8 * pow(1. / hsc, 1. / (double)(s.p + 1));
                s.hScale = hsc;
                if (s.p > MIN_ORDER) {
                    double hm = ep_fact(s.p - 1) * wrms_v(c, s.p);
                    hm = 1. / (hm + EPS_F);
                    hm = 0.8 * pow(hm, 1. / (double)s.p);
                    if (hm > hsc) { s.hScale = hm; s.p -= 1; s.odeOrderStatus = O_DEC; }
                }
                s.hScale = fmin(s.hScale, 0.5);
                s.hScale = fmax(s.hScale, 0. / fabs(s.h));
                s.hScale = fmax(s.hScale, 1.e-3);
                if (s.p == 1 && s.hScale < 0.1 && n.steps > 10 && s.tInMeshPoint != s.tStabilize) {
                    c.rhs(Z0, VA_);
                    n.fcalls++;
                    n.rhs_total++;
                    const double h = s.h;
                    c.each([&](int e) { c.x(Z0 + 1, e) = c.x(VA_, e) * h; });
                    s.tStabilize = s.tInMeshPoint;
                }
                s.odeHStatus = H_DEC;
                s.hNextStep *= s.hScale;
            } else {
                s.hNordsieck = s.h; s.hPreviousStep = s.h; s.hNextStep = s.h;
                s.orderInNextStep = s.p;
                s.odeOrderStatus = O_CONST; s.odeHStatus = H_CONST;
                convergence_rate(c, s, n);
                if (s.iterErrorFailure > 0 || s.iterConvergenceFailure > 0) {
                    s.iterErrorFailure = 0; s.iterConvergenceFailure = 0;
                } else {
                    s.iterOrder++;
                    if (s.iterOrder > (unsigned)s.p + 1) new_order_new_h(c, s, n);
                }
                const int p = s.p;
                c.each([&](int e) {
#pragma unroll
                    for (int i = 0; i <= MAX_ORDER + 2; i++)
                        if (i <= p + 2) c.x(Z0 + i, e) = c.x(V0 + i, e);
                });
                check_constraints(c, Z0, n);
                LSM_BDF_TOC(c, 11, t_tail);
                break;
            }
        }
        LSM_BDF_TOC(c, 11, t_tail);
    }
}

// MultiValueSolver::SetInitialConditions (:368-407) + Solve (:424-617), first call,
// scalar tolerances, min/max constraints on.  Y0_ holds y0 on entry; on success
// Y_ holds the solution and the return value is the OdeStatus (> 0).
template <class C>
__device__ __forceinline__ int bdf_integrate(C &c, const BdfConst &K, double tOut, BdfCounters &n) {
    BdfState s;
    // ResetMethod (MethodGear.hpp:80-124) + Reset (MultiValueSolver.hpp:107-131);
    // fresh solver per cell: alfa2_ = 0.8 and maximum_order_ = 5.
    s.p = 1; s.iterOrder = 0; s.orderInNextStep = 1;
    s.maxIterationsJacobian = 20 + 5 * (c.ne - 3);
    if (s.maxIterationsJacobian > 50 && c.ne < 80) s.maxIterationsJacobian = 50;
    s.odeHStatus = H_CONST; s.odeOrderStatus = O_CONST; s.status = ST_INIT;
    s.factorizationStatus = M_TO_FACTOR; s.jacobianStatus = J_CHANGE; s.convOK = 1;
    s.h = 0.; s.hNextStep = 0.; s.hScale = 0.;
    s.iterConvergence = 0; s.iterConvergenceRate = 0; s.iterConvergenceFailure = 0; s.iterErrorFailure = 0;
    s.stepOfLastJacobian = 0; s.convFailOrderMax = 0;
    s.hPreviousStep = 0.; s.hNordsieck = 0.; s.hScaleMax = 1000.;
    s.fixedOrderMin = 0; s.iterConstSteps = 0; s.iterMinimumOrder = 0;
    s.alfa2_4 = 0.8;
    s.alfa2_5 = 0.8;
    s.maximum_order = MAX_ORDER;
    n = BdfCounters{0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0};
    // Fresh solver per cell: the Nordsieck and work arrays start at zero, as in the oracle, so that
    // nothing of the previous cell or of uninitialised registers reaches this one.  (The reference
    // sizes them once, MethodGear.hpp:180-192, and keeps them from cell to cell.)
    c.each([&](int e) {
#pragma unroll
        for (int i = 0; i < 8; i++) { c.x(Z0 + i, e) = 0.; c.x(V0 + i, e) = 0.; }
        c.x(VB_, e) = 0.; c.x(VA_, e) = 0.; c.x(DB_, e) = 0.; c.x(B_, e) = 0.; c.x(AUX_, e) = 0.; c.x(Y_, e) = 0.;
    });
    s.tOut = K.t0; s.t = K.t0; s.t0 = K.t0; s.tInMeshPoint = K.t0; s.tStabilize = K.t0;
    c.rhs(Y0_, F_);
    n.rhs_total++;
    // Solve
    s.tOut = tOut;
    int illegal = 0;
    c.each([&](int e) {
        const double y = c.x(Y0_, e);
        illegal += c.count(y < c.ymin(e)) + c.count(y > c.ymax(e));
    });
    if (illegal != 0) return ST_ILLEGAL;
    {   // CalculateInitialStepSize, MultiValueSolver.hpp:649-685
        double lo, hi;
        if (s.t0 != 0.) { lo = 100. * EPS_D * fabs(s.t0); hi = fabs(s.t0); }
        else { lo = EPS_D; hi = 1.e-4; }
        one_over_epsilon(c, K, Y0_);
        double h0 = 0.5 / (wrms(c, F_) + EPS_F);
        if (h0 < lo) h0 = lo;
        if (h0 > hi) h0 = hi;
        h0 *= fabs(s.tOut - s.t0) / (s.tOut - s.t0);
        s.first_step = h0;
    }
    c.each([&](int e) {
        c.x(Z0, e) = c.x(Y0_, e);
        c.x(Z0 + 1, e) = c.x(F_, e) * s.first_step;
    });
    s.hPreviousStep = s.hNordsieck = s.hNextStep = s.first_step;
    s.status = ST_CONT;
    for (unsigned k = 1; k <= K.max_steps; k++) {
        one_over_epsilon(c, K, Z0);
        const double tolSafe = EPS_D * wrms(c, Z0);
        if (tolSafe > 1.) return ST_STRICT;
        s.tInMeshPoint = s.t;
        driver_step(c, K, s, n);
        if (s.status <= 0) return s.status;
        n.steps++;
        if (n.max_order < (unsigned)s.p) n.max_order = s.p;
        if (s.tOut == s.t) {
            c.each([&](int e) { c.x(Y_, e) = c.x(Z0, e); });
            return s.status;
        }
        if (s.tOut < s.t) {   // Interpolation, MultiValueSolver.hpp:688-706
            const double hp = (s.tOut - s.t) / s.hNordsieck;
            const int p = s.p;
            c.each([&](int e) {
                double y = c.x(Z0, e);
                if (hp != 0.) {
                    double hr = 1.;
#pragma unroll
                    for (int i = 1; i <= MAX_ORDER; i++)
                        if (i <= p) { hr *= hp; y += c.x(Z0 + i, e) * hr; }
                }
                c.x(Y_, e) = y;
            });
            if (hp != 0.) check_constraints(c, Y_, n);
            c.rhs(Y_, VA_);      // printResults_ = true: the reactor's QR is the final one
            n.rhs_total++;
            return s.status;
        }
    }
    return ST_MAXSTEPS;
}

}  // namespace lsm
