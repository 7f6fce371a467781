// rk4.hip -- RK4 integration of the LLG equation on a periodic L x L lattice (FP64).
//
// Reference: kernel_fusion/llg_sqt_fusion.c
//   euler        :30-93   exchange field h = sum_nn J1 S_nb, f = h x S (v3d.h:48-53)
//   runge_kutta  :98-145  k_s = dt f(stage_s); cur = S + k1/6 + k2/3 + k3/3 + k4/6
//   renormalise  :350-361 S = len * cur / |cur|
// The reference makes ~14 full-array passes per step (9 OpenMP regions); here one
// kernel per step loads a (TR+8) x (TC+8) tile (radius-4 halo: each RK stage widens
// the dependency by one site) into LDS, runs all four stages there, renormalises
// and writes the TR x TC interior once: 24 B read + 24 B written per site-step.
//
// This file is compiled with -ffp-contract=off and keeps the reference's operation
// order (field accumulated neighbour by neighbour, J multiplied per term, v3d_cross
// operand order, the k/6 + k/3 + k/3 + k/6 left-to-right sum, correctly rounded
// divisions), so trajectories equal the FP64 CPU restatement's (IEEE ==); bits can
// differ only in the sign of an exact zero.  Exact shortcuts: the field sums start from
// -0.0 (the additive identity) where the reference starts from 0.0, J == 1 drops the
// multiplies by J, alpha == 0 drops the damping term, stages 1 and 2 use the halved
// increment f dt/2 (k/2 exactly), the quotients k/6, k/3 and the renormalisation take
// FMA-corrected reciprocals, and the square root and division drop their range wrappers
// (|S|^2 stays within 2^+-900 of 1).
#include <map>
#include <mutex>

#include "asd_internal.h"

namespace asd {

LLGParams make_llg(const asd_params &p)
{
    LLGParams q;
    q.J = p.J;
    q.D = p.D;
    q.Bz = p.Bz;
    q.alpha = p.alpha;
    q.dt = p.dt;
    q.len = p.spin_length;
    q.a_over_len = p.alpha / p.spin_length;
    q.den = 1.0 + p.alpha * p.alpha;
    q.inv_den = 1.0 / q.den;
    q.six = 6.0;
    q.three = 3.0;
    q.three_halves = 1.5;
    q.half_dt = 0.5 * p.dt;
    q.two_D = 2.0 * p.D;
    q.w6 = p.dt / 6.0;
    q.w3 = p.dt / 3.0;
    return q;
}

// f(S) at one site given its four neighbours in the reference order
// (row+1, row-1, col-1, col+1; llg_sqt_fusion.c:181-184).
// UNIT_J (J == 1): n * J is n exactly, so the multiplies are dropped.
// UNDAMPED (alpha == 0): (p + 0 q) * 1 is p, so the damping term is dropped; the value
// is the same (IEEE ==), only an exact -0 in p may come out as +0 the other way.
// NO_BZ (B_z == 0): the add of B_z is dropped (hz + 0.0 is hz, up to the sign of a zero).
// FAST (the tolerance mode): the anisotropy term, the cross products and the damping term
// as fused multiply-adds (one rounding where the reference's expressions round twice).
template <bool UNIT_J = false, bool UNDAMPED = false, bool NO_BZ = false, bool FAST = false>
__device__ __forceinline__ void llg_f(const LLGParams &p, double sx, double sy, double sz,
                                      const double *n0, const double *n1, const double *n2,
                                      const double *n3, double &fx, double &fy, double &fz)
{
    // the sums start from -0.0, the additive identity: -0.0 + x is x for every x, so the
    // first add folds away.  The reference starts from 0.0; 0.0 + x differs from x only for
    // x = -0.0, so the sums are equal (IEEE ==) and can differ only in the sign of an exact
    // zero, which no later operation turns into a different value.
    double hx = -0.0, hy = -0.0, hz = -0.0;
    if constexpr (FAST && !UNIT_J) {
        // the tolerance mode: J (n0 + n1 + n2 + n3), one multiply per component
        hx = (((n0[0] + n1[0]) + n2[0]) + n3[0]) * p.J;
        hy = (((n0[1] + n1[1]) + n2[1]) + n3[1]) * p.J;
        hz = (((n0[2] + n1[2]) + n2[2]) + n3[2]) * p.J;
    } else if constexpr (UNIT_J) {
        hx += n0[0]; hy += n0[1]; hz += n0[2];
        hx += n1[0]; hy += n1[1]; hz += n1[2];
        hx += n2[0]; hy += n2[1]; hz += n2[2];
        hx += n3[0]; hy += n3[1]; hz += n3[2];
    } else {
        hx += n0[0] * p.J; hy += n0[1] * p.J; hz += n0[2] * p.J;
        hx += n1[0] * p.J; hy += n1[1] * p.J; hz += n1[2] * p.J;
        hx += n2[0] * p.J; hy += n2[1] * p.J; hz += n2[2] * p.J;
        hx += n3[0] * p.J; hy += n3[1] * p.J; hz += n3[2] * p.J;
    }
    if constexpr (FAST) {
        hz = __builtin_fma(p.two_D, sz, hz);
        if constexpr (!NO_BZ) hz += p.Bz;
        const double px = __builtin_fma(hy, sz, -(hz * sy));
        const double py = __builtin_fma(hz, sx, -(hx * sz));
        const double pz = __builtin_fma(hx, sy, -(hy * sx));
        if constexpr (UNDAMPED) {
            fx = px; fy = py; fz = pz;
        } else {
            const double qx = __builtin_fma(sy, pz, -(sz * py));
            const double qy = __builtin_fma(sz, px, -(sx * pz));
            const double qz = __builtin_fma(sx, py, -(sy * px));
            fx = __builtin_fma(p.a_over_len, qx, px) * p.inv_den;
            fy = __builtin_fma(p.a_over_len, qy, py) * p.inv_den;
            fz = __builtin_fma(p.a_over_len, qz, pz) * p.inv_den;
        }
        return;
    }
    hz += 2.0 * p.D * sz;
    if constexpr (!NO_BZ) hz += p.Bz;
    const double px = hy * sz - hz * sy;
    const double py = hz * sx - hx * sz;
    const double pz = hx * sy - hy * sx;
    if constexpr (UNDAMPED) {
        fx = px; fy = py; fz = pz;
    } else {
        const double qx = sy * pz - sz * py;
        const double qy = sz * px - sx * pz;
        const double qz = sx * py - sy * px;
        fx = (px + p.a_over_len * qx) * p.inv_den;
        fy = (py + p.a_over_len * qy) * p.inv_den;
        fz = (pz + p.a_over_len * qz) * p.inv_den;
    }
}

// a + k / d, with k / d correctly rounded by the FMA correction of Markstein's theorem
// (z = RN(1/d) within half an ulp, q = RN(k z) within one ulp of k/d, r = k - d q exact,
// so RN(q + r z) = RN(k/d)).  Written as -(fma(-r, z, -q)), whose value is RN(q + r z)
// for k != 0 and which keeps the sign of k = +-0; the negation folds into the add.
// Three instructions instead of the IEEE division sequence; same bits as k / d
// (checked against division on 2.2e9 random and near-multiple operands).
__device__ __forceinline__ double add_div(double a, double k, double d, double z)
{
    const double q = k * z;
    const double r = __builtin_fma(-q, d, k);
    return a - __builtin_fma(-r, z, -q);
}

// (len * a) / nrm for the three components with one IEEE division (y = 1 / nrm) and the
// same correction per component: bit-identical to three divisions.
__device__ __forceinline__ void renorm(double len, double ax, double ay, double az, double &ox, double &oy,
                                       double &oz)
{
    const double nrm = sqrt(ax * ax + ay * ay + az * az);
    const double y = 1.0 / nrm;
    auto one = [&](double a) {
        const double n = len * a, q = n * y;
        const double r = __builtin_fma(-q, nrm, n);
        return -__builtin_fma(-r, y, -q);
    };
    ox = one(ax); oy = one(ay); oz = one(az);
}

// The same three values (len a) / sqrt(a.a) for operands in the normal range, in fewer
// instructions.  gfx950's correctly rounded f64 square root and division are Newton
// sequences around v_rsq_f64 / v_rcp_f64, wrapped in operand scaling and special-case
// selection for zeros, infinities and exponents beyond about +-900; here |a|^2 is near
// len^2 (the RK4 increment is a small rotation), so the wrappers are identities and are
// left out:
//   sqrt(x): y = rsq(x); g = x y; h = y / 2; r = 1/2 - h g; g += g r; h += h r;
//            twice: d = x - g g; g += d h                                (as the compiler's)
//   n / b:   y0 = rcp(b); twice: e = 1 - b y; y += y e;   q = n y; r = n - b q;  q + r y
// The reciprocal iterate y depends on b alone, so one serves all three quotients, each
// of which is then exactly the compiler's IEEE division.  Valid while |a|^2 and every
// nonzero len a_i stay above 2^-900 and below 2^900 (the launcher's parameter range
// keeps spin_length there).
// UNIT_LEN (spin_length == 1): the multiply 1 * a (exact) is dropped.
// Returns |a|^2, which the numeric guard checks.
template <bool UNIT_LEN = false>
__device__ __forceinline__ double renorm_fast(double len, double ax, double ay, double az, double &ox, double &oy,
                                              double &oz)
{
    const double x = ax * ax + ay * ay + az * az;
    const double rs = __builtin_amdgcn_rsq(x);
    double g = x * rs, h = rs * 0.5;
    const double r0 = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r0, g);
    h = __builtin_fma(h, r0, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    const double nrm = __builtin_fma(d, h, g);
    double y = __builtin_amdgcn_rcp(nrm);
    double e = __builtin_fma(-nrm, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-nrm, y, 1.0);
    y = __builtin_fma(y, e, y);
    auto one = [&](double a) {
        const double n = UNIT_LEN ? a : len * a, q = n * y;
        const double r = __builtin_fma(-nrm, q, n);
        return __builtin_fma(r, y, q);
    };
    ox = one(ax); oy = one(ay); oz = one(az);
    return x;
}
// The tolerance mode's renormalisation: y = 1/sqrt(|a|^2) from v_rsq_f64 refined by two
// Newton steps y <- y + (y/2)(1 - |a|^2 y^2) (each squares the relative error; the result
// is within a few ulp of 1/|a|), then len * y * a: |S| = len to ~1e-16 in place of the
// correctly rounded quotients.  Returns |a|^2 for the numeric guard.
template <bool UNIT_LEN = false>
__device__ __forceinline__ double renorm_lean(double len, double ax, double ay, double az, double &ox, double &oy,
                                              double &oz)
{
    const double x = __builtin_fma(az, az, __builtin_fma(ay, ay, ax * ax));
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const double r = __builtin_fma(-(x * y), y, 1.0);
        y = __builtin_fma(0.5 * y, r, y);
    }
    const double s = UNIT_LEN ? y : len * y;
    ox = ax * s; oy = ay * s; oz = az * s;
    return x;
}
constexpr double kSixth = 1.0 / 6.0, kThird = 1.0 / 3.0, kTwoThirds = 2.0 / 3.0;

// the numeric guard (Guard in asd_internal.h): |c|^2 outside [2^-900, 2^900] or NaN
__device__ __forceinline__ bool out_of_range(double x) { return !(x >= 0x1p-900 && x <= 0x1p900); }
__device__ __forceinline__ void guard_report(const Guard &g, int traj, int later = 0)
{
    // `later`: the step's offset from the launch's first step (fused steps)
    const long s = (g.ctr ? g.step + (long)g.every * *g.ctr : g.step) + later;
    atomicMin(g.first, ((unsigned long long)s << 20) | (unsigned)traj);
}

__device__ __forceinline__ int wrap(int x, int L)
{
    x %= L;
    return x < 0 ? x + L : x;
}

// NS full RK4 steps (+ renormalisation) of one TR x TC output tile.  The tile and a
// radius-4 NS halo are loaded once; every RK stage widens the dependency by one site, so
// stage s of step m is evaluated on the sites at depth >= 4 m + s (depth: distance from
// the edge of the loaded tile), and step m produces new states for depth >= 4 m + 4 (the
// last step: the tile itself).  Thread t owns the sites t, t + NT, .. of the widest stage
// region (depth >= 1) for all stages and steps, and keeps their states, stage inputs and
// partial sums S + k1/6 + k2/3 + k3/3 + k4/6 (accumulated left to right) in registers.
// LDS holds two arrays of stage inputs: the owner of a site forms its next stage input
// Y = S + kh (or S + k3) once and stores it into the array the current stage does not
// read, so a stage reads the four neighbours' Y (12 doubles a site; forming Y from S and
// the increment on read would take 30), writes one Y and takes one barrier.  The states
// start in the second array (the stage-1 inputs), which stage 2 then overwrites.
// Stages 1 and 2 use the halved increment kh = f (dt/2), which is k/2 exactly
// (RN(f dt)/2 = RN(f dt/2) while k is above the subnormal range), so Y = S + k/2 is
// S + kh and k1/6, k2/3 are the correctly rounded kh/3, kh/1.5; the renormalisation is
// renorm_fast (|S|^2 within 2^+-900 of 1, kept by the parameter check).  Every site goes
// through the same operations whatever the tile shape and NS, so fusing steps changes no
// bit.  Fusing pays where one step is latency-bound: two steps load and store once.
template <int TR, int TC, int NS = 1> struct TileLDS {
    static constexpr int R = 4 * NS, RR = TR + 2 * R, RC = TC + 2 * R, NR = RR * RC;
    double Y[2][3][NR];
};

// STORE (one step, exact arithmetic; the adjoint's recomputation): each output site's stage
// inputs Y2 = S + k1/2, Y3 = S + k2/2, Y4 = S + k3 and its unnormalised sum c = S + k1/6 + k2/3 +
// k3/3 + k4/6 are also written to so.Y2 / Y3 / Y4 / c ([site][xyz] of the lattice).
struct StageOut {
    double *Y2 = nullptr, *Y3 = nullptr, *Y4 = nullptr, *c = nullptr;
};

template <int TR, int TC, bool UJ = false, bool UD = false, int NT = TR * TC, int NS = 1, bool FAST = false,
          bool STORE = false>
__device__ __forceinline__ void rk4_tile(TileLDS<TR, TC, NS> &sm, const LLGParams &p, int L, int tile_r, int tile_c,
                                         const double *__restrict__ in, double *__restrict__ out,
                                         double *__restrict__ snap, const Guard &gd, int traj,
                                         const StageOut &so = StageOut())
{
    constexpr int R = 4 * NS, RR = TR + 2 * R, RC = TC + 2 * R, NR = RR * RC;
    constexpr int W = RC - 2, NW = (RR - 2) * W;  // the widest stage region
    constexpr int NE = (NW + NT - 1) / NT;        // its sites per thread
    static_assert(NT >= TR * TC, "the snapshot writer takes one tile site per thread");
    static_assert(!STORE || (NS == 1 && !FAST), "stage values are stored for one exact step");
    double(&sS)[3][NR] = sm.Y[1];
    const int r0 = tile_r * TR - R, c0 = tile_c * TC - R;
    const int t = threadIdx.x;

    // tile coordinates lie in [-R, L + RR); when L >= RR one correction wraps them,
    // which avoids a runtime integer division per loaded site
    const bool near = L >= RR;
    for (int idx = t; idx < NR; idx += NT) {
        int gr = r0 + idx / RC, gc = c0 + idx % RC;
        if (near) {
            gr += gr < 0 ? L : (gr >= L ? -L : 0);
            gc += gc < 0 ? L : (gc >= L ? -L : 0);
        } else {
            gr = wrap(gr, L);
            gc = wrap(gc, L);
        }
        const double *s = in + 3 * ((long)gr * L + gc);
        sS[0][idx] = s[0]; sS[1][idx] = s[1]; sS[2][idx] = s[2];
    }
    __syncthreads();
    if (snap && t < TR * TC) {  // coalesced: thread t writes tile site t
        const int ir = t / TC, ic = t % TC, grow = tile_r * TR + ir, gcol = tile_c * TC + ic;
        if (grow < L && gcol < L) {
            const int own = (ir + R) * RC + (ic + R);
            double *o = snap + 3 * ((long)grow * L + gcol);
            o[0] = sS[0][own]; o[1] = sS[1][own]; o[2] = sS[2][own];
        }
    }

    // lattice site of this thread's j-th site when it is one of the tile's outputs, else -1
    auto tile_site = [&](int j) -> long {
        const int e = t + j * NT, er = 1 + e / W, ec = 1 + e % W;
        const int grow = tile_r * TR + er - R, gcol = tile_c * TC + ec - R;
        if (e >= NW || er < R || er >= R + TR || ec < R || ec >= R + TC || grow >= L || gcol >= L) return -1;
        return (long)grow * L + gcol;
    };
    // this thread's sites: LDS index, depth, state
    int idx[NE], depth[NE];
    double S[NE][3];
#pragma unroll
    for (int j = 0; j < NE; j++) {
        const int e = t + j * NT;
        const bool has = e < NW;
        const int er = has ? 1 + e / W : 1, ec = has ? 1 + e % W : 1;
        idx[j] = er * RC + ec;
        depth[j] = has ? min(min(er, RR - 1 - er), min(ec, RC - 1 - ec)) : -1;
        S[j][0] = sS[0][idx[j]]; S[j][1] = sS[1][idx[j]]; S[j][2] = sS[2][idx[j]];
    }

#pragma unroll
    for (int m = 0; m < NS; m++) {
        double A[NE][3], y[NE][3];  // partial sums, stage inputs
#pragma unroll
        for (int j = 0; j < NE; j++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                A[j][d] = 0.0;
                y[j][d] = S[j][d];
            }
#pragma unroll
        for (int stage = 1; stage <= 4; stage++) {
            // stage 1 reads the states (array 1), stage s > 1 the inputs stage s - 1 wrote
            const double(&Yin)[3][NR] = sm.Y[stage & 1];
            double(&Yout)[3][NR] = sm.Y[(stage - 1) & 1];
#pragma unroll
            for (int j = 0; j < NE; j++) {
                if (depth[j] < 4 * m + stage) continue;
                const int i = idx[j];
                double n[4][3];
                const int nb[4] = {i + RC, i - RC, i - 1, i + 1};  // row+1, row-1, col-1, col+1
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    n[q][0] = Yin[0][nb[q]]; n[q][1] = Yin[1][nb[q]]; n[q][2] = Yin[2][nb[q]];
                }
                double fx, fy, fz;
                llg_f<UJ, UD, false, FAST>(p, y[j][0], y[j][1], y[j][2], n[0], n[1], n[2], n[3], fx, fy, fz);
                const double h = stage <= 2 ? p.half_dt : p.dt;
                if constexpr (FAST) {
                    // Y = S + f h and the partial sum S + f dt/6 + f dt/3 + ... as single FMAs
                    const double f[3] = {fx, fy, fz};
                    const double wt = (stage == 1 || stage == 4) ? p.w6 : p.w3;
                    if (stage < 4) {
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            y[j][d] = __builtin_fma(f[d], h, S[j][d]);
                            Yout[d][i] = y[j][d];
                        }
                    }
                    if (depth[j] >= 4 * m + 4) {
#pragma unroll
                        for (int d = 0; d < 3; d++) A[j][d] = __builtin_fma(f[d], wt, stage == 1 ? S[j][d] : A[j][d]);
                    }
                    continue;
                }
                const double k[3] = {fx * h, fy * h, fz * h};
                if (stage < 4) {  // the next stage's input: S + kh (stages 1, 2) or S + k3
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        y[j][d] = S[j][d] + k[d];
                        Yout[d][i] = y[j][d];
                    }
                    if constexpr (STORE) {
                        const long gs = tile_site(j);
                        if (gs >= 0) {
                            double *o = (stage == 1 ? so.Y2 : stage == 2 ? so.Y3 : so.Y4) + 3 * gs;
                            o[0] = y[j][0]; o[1] = y[j][1]; o[2] = y[j][2];
                        }
                    }
                }
                if (depth[j] >= 4 * m + 4) {
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        if (stage == 1)
                            A[j][d] = add_div(S[j][d], k[d], p.three, kThird);  // kh/3 = k1/6
                        else if (stage == 2)
                            A[j][d] = add_div(A[j][d], k[d], p.three_halves, kTwoThirds);  // kh/1.5 = k2/3
                        else if (stage == 3)
                            A[j][d] = add_div(A[j][d], k[d], p.three, kThird);
                        else
                            A[j][d] = add_div(A[j][d], k[d], p.six, kSixth);
                    }
                }
            }
            // the last stage of the last step reads nothing from LDS afterwards
            if (stage < 4 || m + 1 < NS) __syncthreads();
        }
        if (m + 1 < NS) {
            // the next step's states, into array 1 (last read by stage 3)
#pragma unroll
            for (int j = 0; j < NE; j++) {
                if (depth[j] < 4 * m + 4) continue;
                const double x = FAST ? renorm_lean(p.len, A[j][0], A[j][1], A[j][2], S[j][0], S[j][1], S[j][2])
                                      : renorm_fast(p.len, A[j][0], A[j][1], A[j][2], S[j][0], S[j][1], S[j][2]);
                if (gd.first && out_of_range(x)) guard_report(gd, traj, m);
                sS[0][idx[j]] = S[j][0]; sS[1][idx[j]] = S[j][1]; sS[2][idx[j]] = S[j][2];
            }
            __syncthreads();
        } else {
#pragma unroll
            for (int j = 0; j < NE; j++) {
                if (depth[j] < R) continue;
                const int e = t + j * NT, er = 1 + e / W, ec = 1 + e % W;
                const int grow = tile_r * TR + er - R, gcol = tile_c * TC + ec - R;
                if (grow >= L || gcol >= L) continue;
                double *o = out + 3 * ((long)grow * L + gcol);
                if constexpr (STORE) {
                    double *oc = so.c + 3 * ((long)grow * L + gcol);
                    oc[0] = A[j][0]; oc[1] = A[j][1]; oc[2] = A[j][2];
                }
                const double x = FAST ? renorm_lean(p.len, A[j][0], A[j][1], A[j][2], o[0], o[1], o[2])
                                      : renorm_fast(p.len, A[j][0], A[j][1], A[j][2], o[0], o[1], o[2]);
                if (gd.first && out_of_range(x)) guard_report(gd, traj, m);
            }
        }
    }
}

// The tile of a block, XCD-aware: blocks are dispatched round-robin over the 8 XCDs (block b
// to XCD b % 8), each with its own L2, and a tile's halo comes from its neighbours' previous
// outputs.  Where the tile grid splits into 8 patches (gx even, gy a multiple of 4), XCD x
// takes patch x (2 x 4 patches of gx / 2 x gy / 4 tiles), so most halo sites were written
// through the same L2; otherwise block (x, y) takes tile (y, x).  A bijection either way.
// Used for tiles of 16 x 16 and up (L = 256 / 512 single lattices 2.6% / 5% faster, 16
// replicas of L = 64 4.4%); the 8 x 8 tiles at L = 128 measured 2.7% slower with it.
template <bool XCD>
__device__ __forceinline__ void xcd_tile(int &tr, int &tc)
{
    const int gx = gridDim.x, gy = gridDim.y;
    if (!XCD || gx % 2 || gy % 4) {
        tr = blockIdx.y;
        tc = blockIdx.x;
        return;
    }
    const int lin = blockIdx.x + gx * blockIdx.y, x = lin % 8, m = lin / 8, pw = gx / 2, ph = gy / 4;
    tr = (x / 2) * ph + m / pw;
    tc = (x % 2) * pw + m % pw;
}

// single lattice (nrep replicas along blockIdx.z); snapshot slot from *snap_counter
// when given (graph replay), else slot 0 of snap
template <int TR, int TC, bool UJ, bool UD, int NT = TR * TC, int NS = 1, bool FAST = false>
__global__ __launch_bounds__(NT) void rk4_step_kernel(LLGParams p, int L,
                                                          const double *__restrict__ in,
                                                          double *__restrict__ out,
                                                          double *__restrict__ snap,
                                                          const int *__restrict__ snap_counter,
                                                          long snap_stride, Guard gd)
{
    __shared__ TileLDS<TR, TC, NS> sm;
    const long base = (long)blockIdx.z * L * L * 3;
    double *sn = nullptr;
    if (snap) sn = snap + (snap_counter ? (long)*snap_counter : 0) * snap_stride + base;
    int tr, tc;
    xcd_tile<(TR >= 16)>(tr, tc);
    rk4_tile<TR, TC, UJ, UD, NT, NS, FAST>(sm, p, L, tr, tc, in + base, out + base, sn, gd, blockIdx.z);
}

// ------------------------------------------------------------------ batching
// Uneven-size independent trajectories in one persistent launch per step (or per two
// fused steps): the grid strides over a work list of (trajectory, tile) items; every
// trajectory has its own L, parameters and state/snapshot buffers (BatchTraj table).  The
// tile shape, threads per tile and steps per launch are one of the variants below, chosen
// by the batch autotuner; every variant gives the same bits (rk4_tile's arithmetic does
// not depend on the shape or on NS).
template <int TR, int TC, int NT = TR * TC, int NS = 1, bool FAST = false>
__global__ __launch_bounds__(NT) void rk4_batch_kernel(const BatchTraj *__restrict__ trajs,
                                                       const int2 *__restrict__ items, int nitems, int parity,
                                                       const int *__restrict__ snap_counter, int snap_mul,
                                                       int snap_add, Guard gd)
{
    __shared__ TileLDS<TR, TC, NS> sm;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 w = items[it];
        const BatchTraj &tr = trajs[w.x];
        const int ntc = (tr.L + TC - 1) / TC;
        const double *in = parity ? tr.b : tr.a;
        double *out = parity ? tr.a : tr.b;
        // snapshot slot *snap_counter * snap_mul + snap_add (none where snap_add < 0)
        double *sn = nullptr;
        if (snap_add >= 0 && tr.snaps)
            sn = tr.snaps + ((snap_counter ? (long)*snap_counter * snap_mul : 0L) + snap_add) * 3L * tr.L * tr.L;
        __syncthreads();  // LDS of the previous item fully consumed
        if (tr.p.J == 1.0 && tr.p.alpha == 0.0)
            rk4_tile<TR, TC, true, true, NT, NS, FAST>(sm, tr.p, tr.L, w.y / ntc, w.y % ntc, in, out, sn, gd, w.x);
        else
            rk4_tile<TR, TC, false, false, NT, NS, FAST>(sm, tr.p, tr.L, w.y / ntc, w.y % ntc, in, out, sn, gd, w.x);
    }
}

// ------------------------------------------------- register-streaming step
// For lattices far larger than the caches.  One wave owns 64 consecutive columns of one
// band of R rows: lanes 4..59 are output columns, lanes 0..3 and 60..63 a 4-column halo
// (each RK stage widens the dependency by one site).  The wave streams down the rows:
// at step t it loads row t and evaluates stage 1 at row t-1, stage 2 at row t-2, stage
// 3 at row t-3 and stage 4 at row t-4, so each stage finds its row+1 / row-1 inputs in
// the same lane's registers and its col-1 / col+1 inputs in the neighbouring lanes (DPP
// whole-wave shifts).  No barriers; LDS holds only each lane's own partial sums.  Each
// site is read once and written once (plus the halo: 64/56 columns, (R+8)/R rows).
// Every site gets the values of the reference's expressions (llg_f, Y = S + k/2.0 or
// S + k, k/6.0 + k/3.0 + k/3.0 + k/6.0 left to right, len * a / |a|) through the same
// exact rewrites as rk4_tile (halved increments, FMA-corrected quotients, the square
// root and division without their range wrappers); the results equal the reference's
// (IEEE ==) and differ in bits at most in the sign of an exact zero.
//
// The kernel is bound by FP64 VALU issue, not by HBM: about 236 VALU instructions per
// row step of a wave (48 of them the DPP lane shifts), 4 cycles each, for 56 output sites.
// Moving the lane shifts to the LDS crossbar (ds_bpermute) measured slower (the LDS
// pipe saturates), and so did a persistent grid of one equal share per resident wave
// (its outer segment loop costs the fourth wave per SIMD) and two columns per lane
// (fewer instructions per site, but two waves per SIMD and a shallower prefetch).
constexpr int kStreamHalo = 4, kStreamCols = 64 - 2 * kStreamHalo;

// bound_ctrl: the edge lane (a halo lane) receives 0, so no old value is needed
__device__ __forceinline__ double lane_from_left(double v)  // lane l receives lane l-1
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_from_right(double v)  // lane l receives lane l+1
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// f(c) with row+1 = up, row-1 = dn and the column neighbours from the adjacent lanes
template <bool UJ, bool UD, bool PL = false, bool FAST = false>
__device__ __forceinline__ void stream_f(const LLGParams &p, const double (&c)[3], const double (&up)[3],
                                         const double (&dn)[3], double (&f)[3])
{
    double lf[3], rt[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lf[d] = lane_from_left(c[d]);
        rt[d] = lane_from_right(c[d]);
    }
    llg_f<UJ, UD, PL, FAST>(p, c[0], c[1], c[2], up, dn, lf, rt, f[0], f[1], f[2]);
}
// k = h f(c) (h = dt, or dt/2 for the halved increment)
template <bool UJ, bool UD, bool PL = false>
__device__ __forceinline__ void stream_k(const LLGParams &p, const double (&c)[3], const double (&up)[3],
                                         const double (&dn)[3], double h, double (&k)[3])
{
    stream_f<UJ, UD, PL>(p, c, up, dn, k);
    k[0] *= h; k[1] *= h; k[2] *= h;
}

// grid: 4 waves per block, one work item (replica, band, strip) per wave; R % 4 == 0,
// 8 <= R <= L - 8, L >= 64 (checked by the launcher)
// PL (B_z == 0 and spin_length == 1, the plain workloads): both exact shortcuts above.
// FAST: the tolerance mode (fused multiply-adds for the stage inputs and the partial sums,
// renorm_lean).
template <bool UJ, bool UD, bool PL, bool FAST = false>
__global__ __launch_bounds__(256) void rk4_stream_kernel(LLGParams p, int L, int nrep, int R, int nband,
                                                         int nstrip, const double *__restrict__ in,
                                                         double *__restrict__ out, double *__restrict__ snap,
                                                         const int *__restrict__ snap_counter, long snap_stride,
                                                         Guard gd)
{
    const int lane = threadIdx.x & 63;
    // the wave's index as a scalar: everything derived from it (band, rows, the row of
    // each load and store) is then scalar arithmetic instead of per-lane VALU work
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long item = (long)blockIdx.x * 4 + wave;
    if (item >= (long)nrep * nband * nstrip) return;  // whole wave; the kernel has no barriers
    const int strip = (int)(item % nstrip), band = (int)((item / nstrip) % nband);
    const int rep = (int)(item / ((long)nstrip * nband));
    const long ns = (long)L * L;
    const double *src = in + rep * ns * 3;
    double *dst = out + rep * ns * 3;
    double *sn = nullptr;
    if (snap) sn = snap + (snap_counter ? (long)*snap_counter : 0) * snap_stride + rep * ns * 3;

    const int c = strip * kStreamCols - kStreamHalo + lane;  // in [-4, L + 116)
    const int gc = ((c % L) + L) % L;
    const bool ocol = lane >= kStreamHalo && lane < 64 - kStreamHalo && c < L;
    const int r0 = band * R;                                 // first output row
    const int rows = min(R, L - r0);                         // output rows of this band
    auto row_of = [&](int t) {                               // lattice row loaded at step t
        int y = r0 - kStreamHalo + t;                        // in [-4, L + R + 4) with R <= L - 8
        y += y < 0 ? L : 0;
        y -= y >= L ? L : 0;
        return y;
    };
    // addresses as a scalar row base plus this lane's 32-bit byte offset (3 gc doubles),
    // which the loads and stores take as SGPR base + VGPR offset with no per-lane
    // 64-bit arithmetic
    const unsigned lane_off = 24u * (unsigned)gc;
    auto at = [&](const double *row_base) {
        return (const double *)((const char *)row_base + lane_off);
    };
    auto at_w = [&](double *row_base) { return (double *)((char *)row_base + lane_off); };
    const long row_stride = 3L * L;
    auto load = [&](int t, double (&v)[3]) {
        const double *q = at(src + row_of(t) * row_stride);
        v[0] = q[0]; v[1] = q[1]; v[2] = q[2];
    };

    // B: ring of the rows loaded, row t in slot t % 8 from its load (PF = 4 steps ahead)
    // to its last use as a state at step t + 3.  Eight slots and an eight-step unroll give
    // every row fixed registers for its whole life: with separate prefetch and state rings
    // the compiler moves each loaded row between them at the loop latch, and that move
    // waits for every load in flight.  Y2/Y3/Y4: stage inputs of the rows in flight.
    double B[8][3], Y2[4][3], Y3[4][3], Y4[4][3];
    // the partial sums S + k1/6 + k2/3 + k3/3 of the four rows in flight live in LDS
    // (each lane only touches its own doubles: no barriers), which keeps the register
    // file at three waves per SIMD
    __shared__ double sA[4][4][3][64];  // [wave][ring slot][component][lane]
    double(&A)[4][3][64] = sA[wave];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            B[a + 4][d] = 0.0; Y2[a][d] = 0.0; Y3[a][d] = 0.0; Y4[a][d] = 0.0;
            A[a][d][lane] = 0.0;
        }
    // steps t = 0 .. nstep-1 (output row t-8 at step t), padded to whole groups of eight
    // so the unrolled loop has no branch around a load (a conditional load makes the
    // compiler copy the old value into the load's destination first, which again waits
    // for every load in flight).  Padding steps compute on valid rows and store nothing;
    // loads past the last step repeat row nstep-1.
    const int nstep = rows + 2 * kStreamHalo;
    const int nstep8 = (nstep + 7) & ~7;
    // prefetch PF = 4 rows ahead (with the partial sums in LDS: three waves per SIMD, no
    // scratch; PF = 3 measured 11% slower at 1024 x 1024 x 64, the partial sums in
    // registers another 8%)
    constexpr int PF = 4;
    // issued in row order, as the loop issues them: the wait at the loop head is the
    // worst case over its entries, and rows loaded in another order here would make it
    // wait for the most recent loads
#pragma unroll
    for (int a = 0; a < PF; a++) {
        __builtin_amdgcn_sched_barrier(0);
        load(a, B[a]);
    }
    __builtin_amdgcn_sched_barrier(0);
    for (int t0 = 0; t0 < nstep8; t0 += 8) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int t = t0 + u;
            // row slots of rows t, t-1, t-2, t-3
            const int b0 = u, b1 = (u + 7) & 7, b2 = (u + 6) & 7, b3 = (u + 5) & 7;
            // stage-input slots of rows t, t-1, .., t-5 (t-4 shares the slot of t)
            const int i0 = u & 3, i1 = (u + 3) & 3, i2 = (u + 2) & 3, i3 = (u + 1) & 3;
            const int i4 = i0, i5 = i1;
            // keep each step's instructions within the step: a load hoisted above the last
            // use of its slot's previous row needs another register, and then a move at the
            // loop latch that waits for every load in flight
            __builtin_amdgcn_sched_barrier(0);
            load(min(t + PF, nstep - 1), B[(u + PF) & 7]);  // the slot of row t-4, now free
            const int tr = t - kStreamHalo;                  // band-relative row of row t
            if (sn && ocol && tr >= 0 && tr < rows) {
                double *o = at_w(sn + (r0 + tr) * row_stride);
                o[0] = B[b0][0]; o[1] = B[b0][1]; o[2] = B[b0][2];
            }
            // Stages 1 and 2 form the halved increment kh = f (dt/2) directly: it is k/2
            // exactly (RN(f dt)/2 = RN(f dt/2) while k is above the subnormal range), so
            // Y = S + k/2 is S + kh and k/6, k/3 are the correctly rounded kh/3, kh/1.5.
            double k[3];
            const int orow = t - 2 * kStreamHalo;  // band-relative output row
            if constexpr (FAST) {
                double f[3];
                stream_f<UJ, UD, PL, true>(p, B[b1], B[b0], B[b2], f);  // stage 1 at row t-1
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    A[i1][d][lane] = __builtin_fma(f[d], p.w6, B[b1][d]);
                    Y2[i1][d] = __builtin_fma(f[d], p.half_dt, B[b1][d]);
                }
                stream_f<UJ, UD, PL, true>(p, Y2[i2], Y2[i1], Y2[i3], f);  // stage 2 at row t-2
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    A[i2][d][lane] = __builtin_fma(f[d], p.w3, A[i2][d][lane]);
                    Y3[i2][d] = __builtin_fma(f[d], p.half_dt, B[b2][d]);
                }
                stream_f<UJ, UD, PL, true>(p, Y3[i3], Y3[i2], Y3[i4], f);  // stage 3 at row t-3
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    A[i3][d][lane] = __builtin_fma(f[d], p.w3, A[i3][d][lane]);
                    Y4[i3][d] = __builtin_fma(f[d], p.dt, B[b3][d]);
                }
                stream_f<UJ, UD, PL, true>(p, Y4[i4], Y4[i3], Y4[i5], f);  // stage 4 at row t-4
#pragma unroll
                for (int d = 0; d < 3; d++) A[i4][d][lane] = __builtin_fma(f[d], p.w6, A[i4][d][lane]);
                if (ocol && orow >= 0 && orow < rows) {
                    double *o = at_w(dst + (r0 + orow) * row_stride);
                    const double x =
                        renorm_lean<PL>(p.len, A[i4][0][lane], A[i4][1][lane], A[i4][2][lane], o[0], o[1], o[2]);
                    if (gd.first && out_of_range(x)) guard_report(gd, rep);
                }
                continue;
            }
            // stage 1 at row t-1: Y1 = S
            stream_k<UJ, UD, PL>(p, B[b1], B[b0], B[b2], p.half_dt, k);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                A[i1][d][lane] = add_div(B[b1][d], k[d], p.three, kThird);  // + k1/6
                Y2[i1][d] = B[b1][d] + k[d];                                // S + k1/2
            }
            // stage 2 at row t-2
            stream_k<UJ, UD, PL>(p, Y2[i2], Y2[i1], Y2[i3], p.half_dt, k);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                A[i2][d][lane] = add_div(A[i2][d][lane], k[d], p.three_halves, kTwoThirds);  // + k2/3
                Y3[i2][d] = B[b2][d] + k[d];                                                  // S + k2/2
            }
            // stage 3 at row t-3
            stream_k<UJ, UD, PL>(p, Y3[i3], Y3[i2], Y3[i4], p.dt, k);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                A[i3][d][lane] = add_div(A[i3][d][lane], k[d], p.three, kThird);
                Y4[i3][d] = B[b3][d] + k[d];
            }
            // stage 4 at row t-4, then the new state of that row
            stream_k<UJ, UD, PL>(p, Y4[i4], Y4[i3], Y4[i5], p.dt, k);
#pragma unroll
            for (int d = 0; d < 3; d++) A[i4][d][lane] = add_div(A[i4][d][lane], k[d], p.six, kSixth);
            if (ocol && orow >= 0 && orow < rows) {
                double *o = at_w(dst + (r0 + orow) * row_stride);
                const double x =
                    renorm_fast<PL>(p.len, A[i4][0][lane], A[i4][1][lane], A[i4][2][lane], o[0], o[1], o[2]);
                if (gd.first && out_of_range(x)) guard_report(gd, rep);
            }
        }
    }
}

// Two RK4 steps per pass over HBM (the tolerance arithmetic only).  The wave runs the row
// pipeline of rk4_stream_kernel twice, the second on the first's renormalised output rows
// while they are still in registers: at step t it loads row t, the first step finishes row
// t-4 and the second finishes row t-8, which is stored.  The dependency widens by eight
// sites: 48 output columns per 64-lane wave, bands of R rows read R + 16 rows.  Each site
// gets exactly the arithmetic of two launches of rk4_stream_kernel<..., FAST = true>, so
// the results equal them bit for bit.  Per spin-step this moves about 31 B of HBM instead
// of 54 B, for 1.25x the VALU work of the one-step kernel (which leaves the VALU half idle).
constexpr int kStream2Halo = 8, kStream2Cols = 64 - 2 * kStream2Halo;

// The four stages of one step of the row pipeline (tolerance arithmetic).  X holds the
// step's input states, row r in slot r % NB; t is the newest row in X (a compile-time
// value after unrolling).  Evaluates stage 1 at row t-1 .. stage 4 at row t-4, leaving the
// partial sum S + k1/6 + k2/3 + k3/3 + k4/6 of row t-4 in A[t & 3].  Y2 / Y3 / Y4 hold the
// stage inputs of the rows in flight, row r in slot r & 3.
template <bool UJ, bool UD, bool PL, int NB, class Acc>
__device__ __forceinline__ void stream_stages_fast(const LLGParams &p, const double (&X)[NB][3], int t,
                                                   double (&Y2)[4][3], double (&Y3)[4][3], double (&Y4)[4][3],
                                                   Acc &&A)
{
    const int x0 = t % NB, x1 = (t + NB - 1) % NB, x2 = (t + NB - 2) % NB, x3 = (t + NB - 3) % NB;
    const int i0 = t & 3, i1 = (t + 3) & 3, i2 = (t + 2) & 3, i3 = (t + 1) & 3;
    double f[3];
    stream_f<UJ, UD, PL, true>(p, X[x1], X[x0], X[x2], f);  // stage 1 at row t-1
#pragma unroll
    for (int d = 0; d < 3; d++) {
        A(i1, d) = __builtin_fma(f[d], p.w6, X[x1][d]);
        Y2[i1][d] = __builtin_fma(f[d], p.half_dt, X[x1][d]);
    }
    stream_f<UJ, UD, PL, true>(p, Y2[i2], Y2[i1], Y2[i3], f);  // stage 2 at row t-2
#pragma unroll
    for (int d = 0; d < 3; d++) {
        A(i2, d) = __builtin_fma(f[d], p.w3, A(i2, d));
        Y3[i2][d] = __builtin_fma(f[d], p.half_dt, X[x2][d]);
    }
    stream_f<UJ, UD, PL, true>(p, Y3[i3], Y3[i2], Y3[i0], f);  // stage 3 at row t-3
#pragma unroll
    for (int d = 0; d < 3; d++) {
        A(i3, d) = __builtin_fma(f[d], p.w3, A(i3, d));
        Y4[i3][d] = __builtin_fma(f[d], p.dt, X[x3][d]);
    }
    stream_f<UJ, UD, PL, true>(p, Y4[i0], Y4[i3], Y4[i1], f);  // stage 4 at row t-4
#pragma unroll
    for (int d = 0; d < 3; d++) A(i0, d) = __builtin_fma(f[d], p.w6, A(i0, d));
}

// One segment of a strip: rows [r0, r0 + rows) of the 48 output columns of strip `strip` of
// one lattice (src -> dst, snapshot rows into sn when set), two steps.  The wave loads rows
// r0 - 8 .. r0 + rows + 7 (periodic; rows <= L).  A1 / A2: the wave's LDS partial sums.
template <bool UJ, bool UD, bool PL>
__device__ __forceinline__ void stream2_segment(const LLGParams &p, int L, int strip, int r0, int rows,
                                                const double *__restrict__ src, double *__restrict__ dst,
                                                double *__restrict__ sn, double (&sA2)[4][3][64], int lane,
                                                const Guard &gd, int rep)
{
    const int c = strip * kStream2Cols - kStream2Halo + lane;  // in [-8, L + 104)
    const int gc = ((c % L) + L) % L;
    const bool ocol = lane >= kStream2Halo && lane < 64 - kStream2Halo && c < L;
    auto row_of = [&](int t) {  // lattice row loaded at step t
        int y = r0 - kStream2Halo + t;  // in [-8, L + 8)
        y += y < 0 ? L : 0;
        y -= y >= L ? L : 0;
        return y;
    };
    const unsigned lane_off = 24u * (unsigned)gc;
    auto at = [&](const double *row_base) { return (const double *)((const char *)row_base + lane_off); };
    auto at_w = [&](double *row_base) { return (double *)((char *)row_base + lane_off); };
    const long row_stride = 3L * L;
    auto load = [&](int t, double (&v)[3]) {
        const double *q = at(src + row_of(t) * row_stride);
        v[0] = q[0]; v[1] = q[1]; v[2] = q[2];
    };

    // B: the loaded rows (slot t % 8, PF = 4 ahead, as in rk4_stream_kernel); S1: the first
    // step's output rows (slot r & 3); Y*, Z*: the two steps' stage inputs; A1: the first
    // step's partial sums (slot r & 3); the second step's sit in LDS (sA2, each lane only
    // touching its own doubles: no barriers), which keeps two waves per SIMD
    double B[8][3], S1[4][3], Y2[4][3], Y3[4][3], Y4[4][3], Z2[4][3], Z3[4][3], Z4[4][3], A1[4][3];
    auto a1 = [&](int i, int d) -> double & { return A1[i][d]; };
    auto a2 = [&](int i, int d) -> double & { return sA2[i][d][lane]; };
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            B[a + 4][d] = 0.0; S1[a][d] = 0.0;
            Y2[a][d] = 0.0; Y3[a][d] = 0.0; Y4[a][d] = 0.0;
            Z2[a][d] = 0.0; Z3[a][d] = 0.0; Z4[a][d] = 0.0;
            A1[a][d] = 0.0; sA2[a][d][lane] = 0.0;
        }
    const int nstep = rows + 2 * kStream2Halo;
    const int nstep8 = (nstep + 7) & ~7;
    constexpr int PF = 4;
#pragma unroll
    for (int a = 0; a < PF; a++) {
        __builtin_amdgcn_sched_barrier(0);
        load(a, B[a]);
    }
    __builtin_amdgcn_sched_barrier(0);
    for (int t0 = 0; t0 < nstep8; t0 += 8) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int t = t0 + u;
            __builtin_amdgcn_sched_barrier(0);
            load(min(t + PF, nstep - 1), B[(u + PF) & 7]);
            const int tr = t - kStream2Halo;  // segment-relative row of row t
            if (sn && ocol && tr >= 0 && tr < rows) {
                double *o = at_w(sn + (r0 + tr) * row_stride);
                o[0] = B[u][0]; o[1] = B[u][1]; o[2] = B[u][2];
            }
            // first step: finishes row t-4 into S1[(t-4) & 3] = S1[u & 3]
            stream_stages_fast<UJ, UD, PL, 8>(p, B, u, Y2, Y3, Y4, a1);
            {
                const int i = u & 3;
                const double x = renorm_lean<PL>(p.len, A1[i][0], A1[i][1], A1[i][2], S1[i][0],
                                                 S1[i][1], S1[i][2]);
                const int r1 = t - 4 - kStream2Halo;  // segment-relative row of row t-4
                if (gd.first && ocol && r1 >= 0 && r1 < rows && out_of_range(x)) guard_report(gd, rep);
            }
            // second step on S1 (newest row t-4, in slot u & 3): finishes row t-8
            stream_stages_fast<UJ, UD, PL, 4>(p, S1, u, Z2, Z3, Z4, a2);
            const int orow = t - 8 - kStream2Halo;
            if (ocol && orow >= 0 && orow < rows) {
                const int i = u & 3;
                double *o = at_w(dst + (r0 + orow) * row_stride);
                const double x = renorm_lean<PL>(p.len, sA2[i][0][lane], sA2[i][1][lane], sA2[i][2][lane], o[0],
                                                 o[1], o[2]);
                if (gd.first && out_of_range(x)) guard_report(gd, rep, 1);
            }
        }
    }
}

// Persistent grid (4 waves per block, as many as are resident): the work is the
// nrep * nstrip * L strip-rows, column by column (replica, strip, row), and wave w takes the
// equal share [w share, (w + 1) share) of it, as one segment per column it touches.  Every
// wave then runs the same number of rows (no partial last round of waves), and the 16 halo
// rows are paid per segment (about one per share) instead of per band.  L >= 64 (checked by
// the launcher).  The guard reports the first step's sites as step `step` and the second's
// as `step` + 1, each output site once.
template <bool UJ, bool UD, bool PL>
__global__ __launch_bounds__(256) void rk4_stream2_kernel(LLGParams p, int L, int nstrip, long total, int share,
                                                          const double *__restrict__ in, double *__restrict__ out,
                                                          double *__restrict__ snap,
                                                          const int *__restrict__ snap_counter, long snap_stride,
                                                          Guard gd)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __shared__ double sA[4][4][3][64];  // [wave][ring slot][component][lane]
    const long ns3 = 3L * L * L;
    double *snap_base = snap ? snap + (snap_counter ? (long)*snap_counter : 0) * snap_stride : nullptr;
    long i = ((long)blockIdx.x * 4 + wave) * share;
    const long iend = min(i + share, total);
    while (i < iend) {
        const long col = i / L;
        const int r0 = (int)(i - col * L);
        const int rows = (int)min((long)(L - r0), iend - i);
        const int rep = (int)(col / nstrip), strip = (int)(col - (long)rep * nstrip);
        stream2_segment<UJ, UD, PL>(p, L, strip, r0, rows, in + rep * ns3, out + rep * ns3,
                                    snap_base ? snap_base + rep * ns3 : nullptr, sA[wave], lane, gd, rep);
        i += rows;
    }
}

__global__ void bump_counter_kernel(int *c) { *c += 1; }
__global__ void set_u64_kernel(unsigned long long *p, unsigned long long v) { *p = v; }

// ------------------------------------------------- compat: euler / runge_kutta

// derivative = f(S), one site per thread (llg_sqt_fusion.c:30-93)
__global__ void rhs_kernel(LLGParams p, int L, long nsites_total, const double *__restrict__ S,
                           double *__restrict__ dS, double scale)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nsites_total) return;
    const long ns = (long)L * L;
    const long rep = g / ns;
    const int i = (int)(g - rep * ns);
    const int row = i / L, col = i % L;
    const double *b = S + rep * ns * 3;
    const double *n0 = b + 3 * ((long)wrap(row + 1, L) * L + col);
    const double *n1 = b + 3 * ((long)wrap(row - 1, L) * L + col);
    const double *n2 = b + 3 * ((long)row * L + wrap(col - 1, L));
    const double *n3 = b + 3 * ((long)row * L + wrap(col + 1, L));
    const double *s = b + 3 * (long)i;
    double fx, fy, fz;
    llg_f(p, s[0], s[1], s[2], n0, n1, n2, n3, fx, fy, fz);
    double *o = dS + 3 * g;
    if (scale != 1.0) {  // k *= dt, as runge_kutta does in place
        fx *= scale; fy *= scale; fz *= scale;
    }
    o[0] = fx; o[1] = fy; o[2] = fz;
}

// cur = spin + k * c  (c = 1/2 as k/2.0, or 1)  -- llg_sqt_fusion.c:111, :121, :131
__global__ void axpy_stage_kernel(long n, const double *__restrict__ spin, const double *__restrict__ k,
                                  double *__restrict__ cur, int half)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    cur[j] = half ? spin[j] + k[j] / 2.0 : spin[j] + k[j];
}

// cur = spin + k1/6 + k2/3 + k3/3 + k4/6  -- llg_sqt_fusion.c:141
__global__ void rk4_combine_kernel(long n, const double *__restrict__ spin, const double *__restrict__ k1,
                                   const double *__restrict__ k2, const double *__restrict__ k3,
                                   const double *__restrict__ k4, double *__restrict__ cur)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    cur[j] = spin[j] + k1[j] / 6.0 + k2[j] / 3.0 + k3[j] / 3.0 + k4[j] / 6.0;
}

int rhs_launch(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *S, double *dS)
{
    const long n = (long)nrep * L * L;
    hipLaunchKernelGGL(rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, L, n,
                       S, dS, 1.0);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

int rk4_compat_stages(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *spin,
                      double *cur, double *k1, double *k2, double *k3, double *k4)
{
    const long ns = (long)nrep * L * L, n = ns * 3;
    const dim3 gs((unsigned)((ns + 255) / 256)), ge((unsigned)((n + 255) / 256)), b(256);
    double *ks[4] = {k1, k2, k3, k4};
    // stage 1 reads spin directly (cur_input = spin, :101-103)
    for (int s = 0; s < 4; s++) {
        hipLaunchKernelGGL(rhs_kernel, gs, b, 0, ctx->stream, p, L, ns, s == 0 ? spin : (const double *)cur,
                           ks[s], p.dt);
        if (s < 3)
            hipLaunchKernelGGL(axpy_stage_kernel, ge, b, 0, ctx->stream, n, spin, (const double *)ks[s], cur,
                               s < 2 ? 1 : 0);
    }
    hipLaunchKernelGGL(rk4_combine_kernel, ge, b, 0, ctx->stream, n, spin, (const double *)k1,
                       (const double *)k2, (const double *)k3, (const double *)k4, cur);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

// Stage inputs of one RK4 step for the adjoint: Y2 = S + k1/2, Y3 = S + k2/2,
// Y4 = S + k3 and the unnormalised c = S + k1/6 + k2/3 + k3/3 + k4/6, with the
// same rounding steps as the fused kernel (so recomputed stages are bit-identical).
// The adjoint's stage recomputation in one launch: the one-step exact tile kernel (8 x 8 tiles,
// 256 threads) with the stage inputs and the unnormalised sum of every site stored.  The values
// are the ones the separate RHS / axpy / combine passes of the reference's runge_kutta produce
// (S + k1/2 as S + kh, the partial sums by the FMA-corrected quotients), since rk4_tile's
// arithmetic is that of the reference (IEEE ==).
template <bool UJ, bool UD>
__global__ __launch_bounds__(256) void rk4_stages_kernel(LLGParams p, int L, const double *__restrict__ in,
                                                         double *__restrict__ scratch, StageOut so)
{
    __shared__ TileLDS<8, 8, 1> sm;
    rk4_tile<8, 8, UJ, UD, 256, 1, false, true>(sm, p, L, blockIdx.y, blockIdx.x, in, scratch, nullptr, Guard(), 0,
                                                so);
}

// S -> Y2 = S + k1/2, Y3 = S + k2/2, Y4 = S + k3, c = S + k1/6 + k2/3 + k3/3 + k4/6 (one lattice)
int rk4_stages_store(asd_ctx *ctx, const LLGParams &p, int L, const double *S, double *Y2, double *Y3, double *Y4,
                     double *c, double *scratch)
{
    StageOut so;
    so.Y2 = Y2; so.Y3 = Y3; so.Y4 = Y4; so.c = c;
    const bool uj = p.J == 1.0, ud = p.alpha == 0.0;
    auto k = uj && ud ? rk4_stages_kernel<true, true>
                      : (uj ? rk4_stages_kernel<true, false>
                            : (ud ? rk4_stages_kernel<false, true> : rk4_stages_kernel<false, false>));
    const dim3 grid((unsigned)((L + 7) / 8), (unsigned)((L + 7) / 8));
    hipLaunchKernelGGL(k, grid, dim3(256), 0, ctx->stream, p, L, S, scratch, so);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

// ------------------------------------------------------------- step launcher

// Batch kernel variants (tile edge, threads per tile, steps per launch) the autotuner times.
// The two-step variants also have a one-step kernel of the same shape for launches that may
// not fuse (a single trailing step, odd snapshot spacing).
struct BatchVariant {
    int tile, threads, steps;
    const char *name;
};
static const BatchVariant kBatchVariants[] = {
    {16, 256, 1, "16x16 tiles, 256 threads, one step per launch"},
    {16, 1024, 2, "16x16 tiles, 1024 threads, two steps per launch"},
    {8, 512, 2, "8x8 tiles, 512 threads, two steps per launch"},
    {32, 1024, 2, "32x32 tiles, 1024 threads, two steps per launch"},
    {8, 256, 1, "8x8 tiles, 256 threads, one step per launch"},
    {32, 1024, 1, "32x32 tiles, 1024 threads, one step per launch"},
};
constexpr int kNumBatchVariants = sizeof(kBatchVariants) / sizeof(kBatchVariants[0]);
int rk4_batch_num_variants() { return kNumBatchVariants; }
const char *rk4_batch_variant_name(int v) { return v >= 0 && v < kNumBatchVariants ? kBatchVariants[v].name : "invalid"; }
int rk4_batch_variant_tile(int v) { return kBatchVariants[v].tile; }
int rk4_batch_variant_steps(int v) { return kBatchVariants[v].steps; }

using BatchKernel = void (*)(const BatchTraj *, const int2 *, int, int, const int *, int, int, Guard);

template <bool F>
static BatchKernel batch_kernel(int v, int ns)
{
    switch (v) {
    case 1: return ns == 2 ? rk4_batch_kernel<16, 16, 1024, 2, F> : rk4_batch_kernel<16, 16, 1024, 1, F>;
    case 2: return ns == 2 ? rk4_batch_kernel<8, 8, 512, 2, F> : rk4_batch_kernel<8, 8, 512, 1, F>;
    case 3: return ns == 2 ? rk4_batch_kernel<32, 32, 1024, 2, F> : rk4_batch_kernel<32, 32, 1024, 1, F>;
    case 4: return rk4_batch_kernel<8, 8, 256, 1, F>;
    case 5: return rk4_batch_kernel<32, 32, 1024, 1, F>;
    default: return rk4_batch_kernel<16, 16, 256, 1, F>;
    }
}

// persistent grid: the blocks of kernel k resident at once (cached per kernel)
// resident blocks per CU of a kernel at `threads` threads (cached per kernel)
static int resident_blocks(const void *k, int threads)
{
    static std::mutex mu;
    static std::map<const void *, int> per_cu;
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_cu.find(k);
    if (it == per_cu.end()) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, threads, 0) != hipSuccess || b < 1) b = 1;
        it = per_cu.emplace(k, b).first;
    }
    return it->second;
}

static int batch_grid(BatchKernel k, int threads, int nitems, int num_cu)
{
    const int n = resident_blocks((const void *)k, threads);
    return std::max(1, std::min(nitems, n * std::max(num_cu, 1)));
}

int rk4_batch_launch(asd_ctx *ctx, int variant, int steps, const BatchTraj *d_trajs, const int2 *d_items,
                     int nitems, int parity, const int *snap_counter, int snap_mul, int snap_add,
                     int bump_counter, const Guard &guard)
{
    ASD_CHECK(variant >= 0 && variant < kNumBatchVariants, "batch variant %d out of range", variant);
    ASD_CHECK(steps == 1 || (steps == 2 && kBatchVariants[variant].steps == 2), "variant %d cannot fuse %d steps",
              variant, steps);
    const BatchKernel k = ctx->rk4_exact ? batch_kernel<false>(variant, steps) : batch_kernel<true>(variant, steps);
    const int threads = kBatchVariants[variant].threads;
    hipLaunchKernelGGL(k, dim3(batch_grid(k, threads, nitems, ctx->num_cu)), dim3(threads), 0, ctx->stream, d_trajs,
                       d_items, nitems, parity, snap_counter, snap_mul, snap_add, guard);
    if (bump_counter)
        hipLaunchKernelGGL(bump_counter_kernel, dim3(1), dim3(1), 0, ctx->stream, (int *)snap_counter);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

template <int TR, int TC, int NT, int NS, bool F>
static auto tile_kernel(bool uj, bool ud)
{
    return uj && ud ? rk4_step_kernel<TR, TC, true, true, NT, NS, F>
                    : (uj ? rk4_step_kernel<TR, TC, true, false, NT, NS, F>
                          : (ud ? rk4_step_kernel<TR, TC, false, true, NT, NS, F>
                                : rk4_step_kernel<TR, TC, false, false, NT, NS, F>));
}

template <int TR, int TC, int NT, int NS = 1>
static void launch_tiles(asd_ctx *ctx, const LLGParams &p, bool uj, bool ud, int L, int nrep, const double *in,
                         double *out, double *snap_out, const int *snap_counter, long snap_stride, const Guard &g)
{
    const dim3 grid((unsigned)((L + TC - 1) / TC), (unsigned)((L + TR - 1) / TR), (unsigned)nrep);
    auto k = ctx->rk4_exact ? tile_kernel<TR, TC, NT, NS, false>(uj, ud) : tile_kernel<TR, TC, NT, NS, true>(uj, ud);
    hipLaunchKernelGGL(k, grid, dim3(NT), 0, ctx->stream, p, L, in, out, snap_out, snap_counter, snap_stride, g);
}

// rows per band of the streaming kernel: 104 (measured at L = 1024 with 64 / 32 / 8
// replicas against 56-248: 781 / 401 / 132 us per step, 120 rows 800 / 428 / 144), or 56
// where 104-row bands would not give 16 waves per CU (8 replicas: 121 us); more, shorter
// bands balance the CUs better at the cost of 8 halo rows per band
static int stream_band_rows(const asd_ctx *ctx, int L, int nrep)
{
    if (L < 64) return 0;  // no streaming below the 64-column wave
    const int nstrip = (L + kStreamCols - 1) / kStreamCols;
    int R = std::min(104, L - 8);
    if ((long)nrep * ((L + R - 1) / R) * nstrip < 16L * std::max(ctx->num_cu, 1)) R = std::min(56, L - 8);
    return R;
}

// the automatic choice among the kernels below: streaming strips once there are at least
// four waves per CU of strips (L >= 64), else LDS tiles by occupancy
static bool rk4_streams(const asd_ctx *ctx, int L, int nrep)
{
    const int R = stream_band_rows(ctx, L, nrep);
    const int nband = L >= 64 ? (L + R - 1) / R : 0, nstrip = (L + kStreamCols - 1) / kStreamCols;
    const long waves = (long)nrep * nband * nstrip;
    return L >= 64 && (ctx->rk4_tile == 64 || (ctx->rk4_tile == 0 && waves >= 4L * ctx->num_cu));
}
static bool rk4_small_tiles(const asd_ctx *ctx, int L, int nrep)
{
    const long n16 = (long)nrep * ((L + 15) / 16) * ((L + 15) / 16);
    return ctx->rk4_tile == 8 || (ctx->rk4_tile == 0 && n16 < std::max(ctx->num_cu, 1));
}

// the automatic choice's 32 x 32 tiles: from half a tile to two tiles per CU
static bool rk4_big_tiles(const asd_ctx *ctx, int L, int nrep)
{
    const long n32 = (long)nrep * ((L + 31) / 32) * ((L + 31) / 32);
    const int cus = std::max(ctx->num_cu, 1);
    return ctx->rk4_tile == 0 && !rk4_small_tiles(ctx, L, nrep) && 2 * n32 > cus && n32 <= 2L * cus;
}
// the automatic choice's 16 x 16 tiles with 1024 threads (four threads per site): below four
// workgroups per CU, where 32 x 32 tiles are not taken
static bool rk4_mid_tiles(const asd_ctx *ctx, int L, int nrep)
{
    const long n16 = (long)nrep * ((L + 15) / 16) * ((L + 15) / 16);
    return ctx->rk4_tile == 0 && !rk4_small_tiles(ctx, L, nrep) && !rk4_big_tiles(ctx, L, nrep) &&
           n16 < 4L * std::max(ctx->num_cu, 1);
}

// automatic choice only, on the 8 x 8, 32 x 32 and 16 x 16 / 1024-thread paths (with
// snapshots every 10: L = 128 6.0 against 7.2 ms per 2000 steps; L = 256 / 320 / 352 8.75 /
// 14.1 / 14.4 against 10.4 / 14.8 / 15.3 ms; 32 x 32 at L = 384 / 448 / 512 / 640 7.8 / 8.25 /
// 7.9 / 14.15 against 8.26 / 8.8 / 8.6 / 14.65 ms per 1000 steps); not for forced tiles, nor
// for 16 x 16 tiles with 256 threads, where the fused halo costs more than the launch it
// saves
// Streaming strips fuse two steps in the tolerance arithmetic only (rk4_stream2_kernel:
// the exact arithmetic keeps the VALU busy already at one step per pass).
int rk4_fusable_steps(const asd_ctx *ctx, int L, int nrep)
{
    if (ctx->rk4_tile != 0) return 1;
    if (rk4_streams(ctx, L, nrep)) return ctx->rk4_exact ? 1 : 2;
    return rk4_small_tiles(ctx, L, nrep) || rk4_big_tiles(ctx, L, nrep) || rk4_mid_tiles(ctx, L, nrep) ? 2 : 1;
}

template <bool F>
static auto stream_kernel(bool uj, bool ud, bool pl)
{
    return uj && ud ? (pl ? rk4_stream_kernel<true, true, true, F> : rk4_stream_kernel<true, true, false, F>)
                    : (uj ? rk4_stream_kernel<true, false, false, F>
                          : (ud ? rk4_stream_kernel<false, true, false, F> : rk4_stream_kernel<false, false, false, F>));
}

int rk4_launch_step(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *in, double *out,
                    double *snap_out, const int *snap_counter, long snap_stride, int bump_counter, int steps,
                    const Guard &guard)
{
    ASD_CHECK(steps == 1 || (steps == 2 && rk4_fusable_steps(ctx, L, nrep) == 2), "cannot fuse %d steps here",
              steps);
    // LDS tiles: 16 x 16 (256 threads, 27.6 KB LDS; 2.25 sites loaded and 5.7
    // f-evaluations per output site) or 32 x 32 (1024 threads; one step: 76.8 KB, two
    // workgroups per CU, 1.56 loads and 4.8 f-evaluations per site; two fused steps: 110.6 KB,
    // one workgroup per CU); the automatic choice below
    const bool big = ctx->rk4_tile == 32;
    // J == 1 and alpha == 0 (the plain workloads) take the specialised arithmetic
    const bool uj = p.J == 1.0, ud = p.alpha == 0.0;
    // streaming strips: forced (64), or automatic once there are at least four waves per
    // CU (each wave's rows form one serial chain, so small lattices are faster in tiles;
    // measured: L = 1024 x 8 replicas 205 us vs 275 us per step, x 1 replica 114 vs 41)
    const int R = stream_band_rows(ctx, L, nrep);
    const int nband = L >= 64 ? (L + R - 1) / R : 0, nstrip = (L + kStreamCols - 1) / kStreamCols;
    const long waves = (long)nrep * nband * nstrip;
    if (rk4_streams(ctx, L, nrep) && steps == 2) {
        const bool pl = p.Bz == 0.0 && p.len == 1.0;
        ASD_CHECK(L >= 64, "two-step strips need L >= 64");
        auto k = uj && ud ? (pl ? rk4_stream2_kernel<true, true, true> : rk4_stream2_kernel<true, true, false>)
                          : (uj ? rk4_stream2_kernel<true, false, false>
                                : (ud ? rk4_stream2_kernel<false, true, false> : rk4_stream2_kernel<false, false, false>));
        // one equal share of the strip-rows per resident wave (at least 8 rows)
        const int nstrip2 = (L + kStream2Cols - 1) / kStream2Cols;
        const long total = (long)nrep * nstrip2 * L;
        const long slots = 4L * resident_blocks((const void *)k, 256) * std::max(ctx->num_cu, 1);
        const int share = (int)std::max<long>(8, (total + slots - 1) / slots);
        const long nwaves = (total + share - 1) / share;
        hipLaunchKernelGGL(k, dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, ctx->stream, p, L, nstrip2, total,
                           share, in, out, snap_out, snap_counter, snap_stride, guard);
    } else if (rk4_streams(ctx, L, nrep)) {
        const bool pl = p.Bz == 0.0 && p.len == 1.0;
        auto k = ctx->rk4_exact ? stream_kernel<false>(uj, ud, pl) : stream_kernel<true>(uj, ud, pl);
        hipLaunchKernelGGL(k, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, ctx->stream, p, L, nrep, R, nband,
                           nstrip, in, out, snap_out, snap_counter, snap_stride, guard);
    } else if (big) {
        launch_tiles<32, 32, 1024>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
    } else {
        // tiles by occupancy, n16 / n32 tiles of 16 x 16 / 32 x 32 in the launch (per step,
        // one lattice unless stated): 8 x 8 tiles (four threads per output site) while 16 x 16
        // tiles would not give every CU a workgroup (L = 128: 3.6 us, 16 x 16 with 256
        // threads 5.3); 32 x 32 from half a tile to two tiles per CU (L = 384 / 448 / 512 /
        // 640: 7.6 / 8.2 / 8.0 / 14.2 us against 8.4 / 11.4 / 10.4 / 15.5 for 16 x 16; 96 x 16,
        // 192 x 4, 128 x 16, 256 x 4 replicas 17-28% faster; at n32 = CUs / 2, 128 x 8
        // replicas, 16 x 16 with 1024 threads 6.4 us against 7.4); 16 x 16 with four threads
        // per site below four workgroups per CU (L = 256: 4.7 us against 5.8 and 6.9), then
        // 16 x 16 with one (L = 768: 18.4 us against 19.5; L = 1024: 29 us against 27 and
        // 47 for streaming strips)
        // two fused steps: 8 x 8 tiles with 512 threads or 16 x 16 tiles with 1024 (one site of
        // the widest stage region, 22 x 22 or 30 x 30, per thread), or 32 x 32 tiles with 1024
        // (three sites of the 46 x 46 region per thread)
        if (steps == 2 && rk4_big_tiles(ctx, L, nrep))
            launch_tiles<32, 32, 1024, 2>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
        else if (steps == 2 && rk4_mid_tiles(ctx, L, nrep))
            launch_tiles<16, 16, 1024, 2>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
        else if (steps == 2)
            launch_tiles<8, 8, 512, 2>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
        else if (rk4_small_tiles(ctx, L, nrep))
            launch_tiles<8, 8, 256>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
        else if (rk4_big_tiles(ctx, L, nrep))
            launch_tiles<32, 32, 1024>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
        else if (rk4_mid_tiles(ctx, L, nrep))
            launch_tiles<16, 16, 1024>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
        else
            launch_tiles<16, 16, 256>(ctx, p, uj, ud, L, nrep, in, out, snap_out, snap_counter, snap_stride, guard);
    }
    if (bump_counter)
        hipLaunchKernelGGL(bump_counter_kernel, dim3(1), dim3(1), 0, ctx->stream, (int *)snap_counter);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

}  // namespace asd
