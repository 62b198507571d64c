// bt_kernels.hip -- gfx950 kernels of the BaryTree interaction path.
//
// Layout in HBM (all tree-ordered, SoA, FP64):
//   targets  tx,ty,tz,tq[M]         sources sx,sy,sz,sq,sw[N]
//   source clusters cx,cy,cz,cq,cw[numnodes*P]  (grid point j = k1 + nl*k2 + nl^2*k3)
//   target clusters Q,W[numnodes*P]; node tables tab[numnodes*3*nl]
// Work lists are CSR arrays built on the host (bt_solver.cpp).
//
// Kernel map to the reference (src/kernels, src/interaction_compute):
//   k_interact<..., GRID=0>  : K_*_PP and K_*_PC_Lagrange (+SS) for a target leaf
//   k_interact<..., GRID=1>  : K_*_CP_Lagrange, K_*_SS_CP/CC_Lagrange for a target grid
//   k_anterp                 : pc_comp_ms_modifiedF(_SS)(_child_to_parent), clusters.c
//   k_down                   : cp_comp_pot(_SS)(_parent_to_child), downpass.c
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <type_traits>

#include "bt_kernels.h"
#include "bt_rsq.h"

namespace btk {

namespace {

constexpr int TPB = 256;  // threads per block of the interaction kernels (4 waves)
constexpr int GRID_TPT = 2;  // targets per thread of grid-target tiles
constexpr int NQ = TILE_QUEUES;  // tile queues, one per XCD
constexpr int KC = 1, KY = 2, KR = 3;


// ---------------------------------------------------------------------
// e^x in FP64 for x <= 0, with a table of 2^(j/64), j < 64 (EXP_TAB
// doubles, filled by exp_table).  x is clamped at -746, where e^x is below
// half the least subnormal.  x = (64 k + j) ln2/64 + f with t = 64 k + j
// the integer nearest x 64/ln2 and |f| <= ln2/128 (ln2/64 split in a
// 36-bit head, exact times |t| < 2^17, and a tail); e^f by its Taylor
// polynomial to degree 5 (truncation below 3.5e-17 relative); the result
// 2^k tab[j] e^f is within about 2 ulp of e^x.
// ---------------------------------------------------------------------
constexpr int EXP_TAB = 64;

__device__ __forceinline__ void exp_table(double *tab, int tid)
{
    if (tid < EXP_TAB) tab[tid] = exp2((double)tid * (1.0 / EXP_TAB));
}

__device__ __forceinline__ double exp_nonpos(double x, const double *tab)
{
    constexpr double INV = 92.33248261689366;                // 64 / ln 2
    constexpr double HI = 0x1.62e42fefa0000p-7, LO = 2.572804622327669e-14;  // ln2/64
    x = fmax(x, -746.0);
    const double t = __builtin_rint(x * INV);
    const double f = fma(-t, LO, fma(-t, HI, x));
    const int n = (int)t;
    double p = fma(f, 1.0 / 120.0, 1.0 / 24.0);
    p = fma(f, p, 1.0 / 6.0);
    p = fma(f, p, 0.5);
    p = fma(f, p, 1.0);
    p = fma(f, p, 1.0);
    return __builtin_amdgcn_ldexp(tab[n & (EXP_TAB - 1)] * p, n >> 6);
}

// e^x, x <= 0: the table form (TAB) or the library's
template <bool TAB>
__device__ __forceinline__ double exp_np(double x, const double *tab)
{
    if constexpr (TAB) return exp_nonpos(x, tab);
    else return exp(x);
}

// ---------------------------------------------------------------------
// The kernels beside the Coulomb family, FP64, one pair's term with the
// charge a (b, tq: subtraction weight and target charge):
//   regularized Yukawa SK: a e^{-kr}/sqrt(r^2+e^2)   SS: (a - tq b) e^{-kr}/sqrt(r^2+e^2)
//   atan (z, period L)   : a (atan(f tan(pi d))/pi - d), d = (dz/L) wrapped to [-1/2, 1/2];
//                          the sum is scaled by wadj afterwards
//   sin-over-r           : a sin(kr)/r        (particle sources: skip r <= DBL_MIN)
//   mq (z, period L)     : a (d c2 / (2 sqrt(d^2 + c1)) - d)
//   rbs-u, rbs-v (x, y)  : a dy / (2 pi (dx^2+dy^2+delta^2)), -a dx / (2 pi (...))
//   user                 : a e^{-kr}/r        (particle sources: skip r <= DBL_MIN)
// (regularized-yukawa_{pp,ss_pp,pc,ss_pc,cp}.c, atan_{pp,pc}.c,
// sin-over-r_{pp,pc,cp}.c, mq_{pp,pc}.c, rbs-u_*.c, rbs-v_*.c, user_kernel_*.c)
// A division by a square root is a product with rsq64 (bt_rsq.h; about 1 ulp, against
// two correctly rounded operations in the reference's form).  dz/L stays a
// division: atan and MQ jump at d = +-1/2, where the wrap must see the
// reference's rounding of d.
// ---------------------------------------------------------------------
constexpr int KRY = KERNEL_REG_YUKAWA, KAT = KERNEL_ATAN, KSR = KERNEL_SIN_OVER_R, KMQ = KERNEL_MQ,
              KRU = KERNEL_RBS_U, KRV = KERNEL_RBS_V, KUS = KERNEL_USER;

__device__ __forceinline__ double wrap_half(double d)
{
    if (d < -0.5) d += 1.0;
    if (d > 0.5) d -= 1.0;
    return d;
}

template <int KER, bool SS, bool CHECK>
__device__ __forceinline__ double other_term(double dx, double dy, double dz, double r2, double a, double b,
                                             double tq, const KParams &kp)
{
    if constexpr (KER == KRY) {
        const double r = sqrt(r2);
        const double e = exp(-kp.p0 * r) * rsq64(r2 + kp.eps2);
        return (SS ? fma(-tq, b, a) : a) * e;
    } else if constexpr (KER == KAT) {
        const double d = wrap_half(dz / kp.p0);
        return a * (1.0 / M_PI * atan(kp.c1 * tan(M_PI * d)) - d);
    } else if constexpr (KER == KMQ) {
        const double d = wrap_half(dz / kp.p0);
        return a * (0.5 * d * kp.c2 * rsq64(d * d + kp.c1) - d);
    } else if constexpr (KER == KRU || KER == KRV) {
        const double r = fma(dy, dy, dx * dx) + kp.c1;
        return (KER == KRU ? 1.0 : -1.0) / (2.0 * M_PI) * a * (KER == KRU ? dy : dx) / r;
    } else {  // KSR, KUS
        const double r = sqrt(r2), ir = rsq64(r2);
        double v = (KER == KSR ? sin(kp.p0 * r) : exp(-kp.p0 * r)) * ir;
        if constexpr (CHECK) v = r > DBL_MIN ? v : 0.0;
        return a * v;
    }
}

// ---------------------------------------------------------------------
// Pair kernels.  A staged source point carries (a, b):
//   particles: a = q (skipping) or q*w (subtraction), b = w
//   clusters : a = Q, b = W (subtraction weights)
// Target-particle form (adds to the potential):
//   Coulomb   SK: a/r                         SS: (a - tq*b*exp(-r^2/k^2))/r
//   Yukawa    SK: a e^{-kr}/r                 SS: (a - tq*b) e^{-kr}/r
//   RegCoul   SK: a/sqrt(r^2+e^2)             SS: (a - tq*b*exp(-r^2/al^2))/sqrt(r^2+e^2)
// (coulomb_pp.c:48-49, coulomb_ss_pp.c:46-47, coulomb_pc.c:47-50, coulomb_ss_pc.c:47,
//  yukawa_pp.c:44-45, yukawa_ss_pp.c:47-48, yukawa_ss_pc.c:46, regularized-coulomb_*.c)
// CHECK applies the reference's singularity skip (r^2 > DBL_MIN, r > DBL_MIN)
// for Coulomb and Yukawa.  The host clears CHECK for source ranges whose
// bounding box cannot share a point with the targets' box, where the skip
// can never trigger; cluster sources never need it (an accepted cluster is
// separated from its targets).
// ---------------------------------------------------------------------
// TAB: exponentials by exp_nonpos from the table tab.
template <int KER, bool SS, bool CHECK, bool TAB = false>
__device__ __forceinline__ void pair_pot(double dx, double dy, double dz, double a, double b,
                                         double tq, const KParams &kp, double &acc,
                                         const double *tab = nullptr)
{
    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    if constexpr (KER == KC) {
        double ir = rsq64(r2);
        if constexpr (CHECK) ir = r2 > (SS ? 0.0 : DBL_MIN) ? ir : 0.0;
        if constexpr (!SS) acc = fma(a, ir, acc);
        else acc = fma(fma(-tq * b, exp_np<TAB>(-r2 * kp.inv_a2, tab), a), ir, acc);
    } else if constexpr (KER == KY) {
        const double ir = rsq64(r2);
        double e = exp_np<TAB>(-kp.p0 * (r2 * ir), tab) * ir;
        if constexpr (CHECK) e = r2 > 0.0 ? e : 0.0;
        if constexpr (!SS) acc = fma(a, e, acc);
        else acc = fma(fma(-tq, b, a), e, acc);
    } else if constexpr (KER == KR) {
        const double ir = rsq64(r2 + kp.eps2);
        if constexpr (!SS) acc = fma(a, ir, acc);
        else acc = fma(fma(-tq * b, exp_np<TAB>(-r2 * kp.inv_a2, tab), a), ir, acc);
    } else {
        acc += other_term<KER, SS, CHECK>(dx, dy, dz, r2, a, b, tq, kp);
    }
}

// Target-grid form (accumulates Q and, for subtraction, W):
//   Coulomb   SK: Q += a/r          SS: Q += a/r, W += exp(-r^2/k^2) b/r
//   Yukawa    SK: Q += a e^{-kr}/r  SS: Q += a e^{-kr}/r, W += b e^{-kr}/r
//   RegCoul   SK: Q += a/sqrt(r^2+e^2)
// (coulomb_cp.c:46-55, coulomb_ss_cp.c:54-55, coulomb_ss_cc.c:54-55,
//  yukawa_cp.c:52, yukawa_ss_cp.c:55-56, yukawa_ss_cc.c:55-56, regularized-coulomb_cp.c:52)
// The reference applies no skip in these kernels.
// grid_term: the same from r^2 (Coulomb, Yukawa, regularized Coulomb);
// EPS_IN: r2 already holds r^2 + e^2 (regularized Coulomb).
template <int KER, bool SS, bool TAB = false, bool EPS_IN = false>
__device__ __forceinline__ void grid_term(double r2, double a, double b, const KParams &kp, double &q,
                                          double &w, const double *tab = nullptr)
{
    if constexpr (KER == KC) {
        const double ir = rsq64(r2);
        q = fma(a, ir, q);
        if constexpr (SS) w = fma(exp_np<TAB>(-r2 * kp.inv_a2, tab) * b, ir, w);
    } else if constexpr (KER == KY) {
        const double ir = rsq64(r2);
        const double e = exp_np<TAB>(-kp.p0 * (r2 * ir), tab) * ir;
        q = fma(a, e, q);
        if constexpr (SS) w = fma(b, e, w);
    } else {
        static_assert(KER == KR, "grid_term: Coulomb, Yukawa or regularized Coulomb");
        q = fma(a, rsq64(EPS_IN ? r2 : r2 + kp.eps2), q);
    }
}

template <int KER, bool SS, bool TAB = false>
__device__ __forceinline__ void pair_grid(double dx, double dy, double dz, double a, double b,
                                          const KParams &kp, double &q, double &w,
                                          const double *tab = nullptr)
{
    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    if constexpr (KER == KC || KER == KY || KER == KR) {
        grid_term<KER, SS, TAB>(r2, a, b, kp, q, w, tab);
    } else {
        q += other_term<KER, false, false>(dx, dy, dz, r2, a, 0.0, 0.0, kp);
    }
}

// ---------------------------------------------------------------------
// Mixed precision pair kernels.  The displacement and r^2 are formed in FP64
// and r^2 is rounded once to FP32 (so its relative error is 2^-24 even for
// the closest pairs), the kernel value G is evaluated in FP32 with the hardware
// v_rsq_f32 / v_exp_f32, and the product with the FP64 charge is
// accumulated per target in FP64.  Each FP32 term is exact in FP64, so the
// accumulation carries no FP32 rounding (the compensation of a Kahan sum,
// at two instructions per pair instead of four).  For subtraction the
// combination (a - tq*b*g) is FP32.
// kp32: {c_gauss = -log2(e)/k^2, c_yuk = -log2(e)*kappa, eps2}
// ---------------------------------------------------------------------
struct KP32 { float cg, cy, eps2; };

__device__ __forceinline__ float rsq32(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float exp2_32(float x) { return __builtin_amdgcn_exp2f(x); }

template <int KER, bool SS, bool CHECK>
__device__ __forceinline__ void pair_pot32(float r2, double a, float af, float bf, float tqf,
                                           const KP32 &k, double &acc)
{
    if constexpr (KER == KC) {
        float ir = rsq32(r2);
        if constexpr (CHECK) ir = r2 > 0.0f ? ir : 0.0f;
        if constexpr (!SS) acc = fma(a, (double)ir, acc);
        else acc += (double)(fmaf(-tqf * bf, exp2_32(r2 * k.cg), af) * ir);
    } else if constexpr (KER == KY) {
        const float ir = rsq32(r2);
        float e = exp2_32((r2 * ir) * k.cy) * ir;
        if constexpr (CHECK) e = r2 > 0.0f ? e : 0.0f;
        if constexpr (!SS) acc = fma(a, (double)e, acc);
        else acc += (double)(fmaf(-tqf, bf, af) * e);
    } else {
        const float ir = rsq32(r2 + k.eps2);
        if constexpr (!SS) acc = fma(a, (double)ir, acc);
        else acc += (double)(fmaf(-tqf * bf, exp2_32(r2 * k.cg), af) * ir);
    }
}

// Subtraction kernels in the mixed near field with coordinates pre-scaled by
// lam (lam^2 = log2(e)/kappa^2 for Coulomb, log2(e)/alpha^2 for regularized
// Coulomb; lam = kappa*log2(e) for Yukawa) and charges a, b pre-multiplied by
// lam, so that the exponentials take their argument without a multiply:
//   Coulomb : (a - tq b 2^(-r'^2)) / r'          = (a - tq b e^(-r^2/k^2)) / r
//   Yukawa  : (a - tq b) 2^(-r') / r'            = (a - tq b) e^(-k r) / r
//   RegCoul : (a - tq b 2^(-r'^2)) / sqrt(r'^2 + e'^2)
// (r' = lam r, e'^2 = lam^2 eps^2).  The two parts are summed apart, p1 the
// a terms and p2 the b terms, and the target combines them as p1 - tq p2
// (one instruction less per pair than forming the difference per pair).
template <int KER, bool CHECK>
__device__ __forceinline__ void pair_ss32(float r2s, float af, float bf, float eps2s, float &p1, float &p2)
{
    if constexpr (KER == KC) {
        float ir = rsq32(r2s);
        if constexpr (CHECK) ir = r2s > 0.0f ? ir : 0.0f;
        p1 = fmaf(af, ir, p1);
        p2 = fmaf(bf, exp2_32(-r2s) * ir, p2);
    } else if constexpr (KER == KY) {
        const float ir = rsq32(r2s);
        float e = exp2_32(-(r2s * ir)) * ir;
        if constexpr (CHECK) e = r2s > 0.0f ? e : 0.0f;
        p1 = fmaf(af, e, p1);
        p2 = fmaf(bf, e, p2);
    } else {
        const float ir = rsq32(r2s + eps2s);
        p1 = fmaf(af, ir, p1);
        p2 = fmaf(bf, exp2_32(-r2s) * ir, p2);
    }
}

// Grid targets with coordinates and charges scaled by lam as in pair_ss32
// (lam = kappa log2(e) for Yukawa; skipping Coulomb and regularized Coulomb:
// 1): the Q part into p1, the W part (subtraction) into p2.
template <int KER, bool SS>
__device__ __forceinline__ void pair_grid_scaled(float r2s, float af, float bf, float eps2s, float &p1, float &p2)
{
    if constexpr (KER == KC) {
        const float ir = rsq32(r2s);
        p1 = fmaf(af, ir, p1);
        if constexpr (SS) p2 = fmaf(bf, exp2_32(-r2s) * ir, p2);
    } else if constexpr (KER == KY) {
        const float ir = rsq32(r2s);
        const float e = exp2_32(-(r2s * ir)) * ir;
        p1 = fmaf(af, e, p1);
        if constexpr (SS) p2 = fmaf(bf, e, p2);
    } else {
        p1 = fmaf(af, rsq32(r2s + eps2s), p1);
    }
}

template <int KER, bool SS>
__device__ __forceinline__ void pair_grid32(float r2, double a, double b, const KP32 &k, double &q,
                                            double &w)
{
    if constexpr (KER == KC) {
        const float ir = rsq32(r2);
        q = fma(a, (double)ir, q);
        if constexpr (SS) w = fma(b, (double)(exp2_32(r2 * k.cg) * ir), w);
    } else if constexpr (KER == KY) {
        const float ir = rsq32(r2);
        const double e = (double)(exp2_32((r2 * ir) * k.cy) * ir);
        q = fma(a, e, q);
        if constexpr (SS) w = fma(b, e, w);
    } else {
        q = fma(a, (double)rsq32(r2 + k.eps2), q);
    }
}

// ---------------------------------------------------------------------
// Interaction tile.  One block = one tile of up to TPB*TPT targets of one
// key (a target leaf, or the grid of a target node); thread tid owns targets
// tid + u*TPB.  Sources stream through LDS in chunks of TPB points: first the
// key's source clusters, then its source particle ranges.  Each target's sum
// stays in registers and is stored once, so no atomics are needed.  A wave
// evaluates only the target slots u that hold a target for some lane (a
// wave-uniform count), so a partly filled tile costs about its targets.
// MIXED selects the FP32 pair kernels above; everything else (staging,
// targets, totals) is FP64.
// ---------------------------------------------------------------------

template <int KER, bool SS, bool GRID, bool MIXED, int TPT>
__device__ __forceinline__ void interact_tile(const InteractArgs &A, const Tile tile, double4 *Ls,
                                              double *Lb, float2 *Lf, const double *Ltab)
{
    const int key = tile.x;
    const int tid = threadIdx.x;
    // target slots with at least one target in this wave (0..TPT), wave-uniform
    const int wbase = __builtin_amdgcn_readfirstlane(tid & ~63);
    const int nu = max(0, min(TPT, (tile.z - wbase + TPB - 1) / TPB));

    double Tx[TPT], Ty[TPT], Tz[TPT], Tq[TPT], acc[TPT], accw[TPT];
    int tix[TPT];
    // target grid: degree, node tables and output row of this tile
    int nl = A.nl;
    const double *ttab = nullptr;
    size_t orow = 0;
    if constexpr (GRID) {
        if (A.multi) {
            const int dg = tile.w >> DEG_SHIFT;
            nl = dg + 1;
            ttab = A.dtab[dg] + (size_t)(tile.w & SLOT_MASK) * 3 * nl;
        } else {
            ttab = A.tnode_tab + (size_t)tile.w * 3 * nl;
        }
        orow = A.out_row ? (size_t)A.out_row[key] : (size_t)tile.w * A.P;
    }
    const KP32 k32{(float)(-1.4426950408889634 * A.kp.inv_a2), (float)(-1.4426950408889634 * A.kp.p0),
                   (float)A.kp.eps2};
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        const int local = tid + u * TPB;
        const int lc = local < tile.z ? local : 0;
        tix[u] = local < tile.z ? tile.y + local : -1;
        if constexpr (GRID) {
            const int j = tile.y + lc;
            const int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
            Tx[u] = ttab[k1]; Ty[u] = ttab[nl + k2]; Tz[u] = ttab[2 * nl + k3]; Tq[u] = 0.0;
        } else {
            const int i = tile.y + lc;
            Tx[u] = A.tx[i]; Ty[u] = A.ty[i]; Tz[u] = A.tz[i];
            Tq[u] = SS ? A.tq[i] : 0.0;
        }
        acc[u] = 0.0; accw[u] = 0.0;
    }
    // Mixed subtraction kernels on particle targets: coordinates relative to
    // the key's centre c and pre-scaled by lam (pair_ss32), T' = lam (t - c)
    // and S' = lam (s - c); r'^2 = |T'|^2 + |S'|^2 - 2 T'.S' (four FP64
    // instructions; T[xyz] hold -2 T', T2 holds |T'|^2, a staged source's w
    // holds |S'|^2).  Its rounding error is below 1e-15 (|T'|^2 + |S'|^2), so
    // in ranges flagged SEG_CLOSE a block of 16 sources with a pair below the
    // key's threshold (lam^2 |key_ctr.w|, less e'^2 for the regularized
    // kernel) is evaluated again from coordinate differences, with the skip of
    // coincident pairs.  FP32 sums over blocks of 32 sources.
    constexpr bool SCALED = MIXED && SS && !GRID;
    double lam = 1.0, th = 0.0, T2[TPT];
    float eps2s = 0.0f;
    double4 ctr = make_double4(0.0, 0.0, 0.0, 0.0);
    bool can_close = false;
    if constexpr (SCALED) {
        lam = KER == KY ? 1.4426950408889634 * A.kp.p0 : sqrt(1.4426950408889634 * A.kp.inv_a2);
        eps2s = (float)(lam * lam * A.kp.eps2);
        ctr = A.key_ctr[key];
        th = lam * lam * (fabs(ctr.w) - (KER == KR ? A.kp.eps2 : 0.0));
        can_close = th > 0.0;
#pragma unroll
        for (int u = 0; u < TPT; u++) {
            const double x = lam * (Tx[u] - ctr.x), y = lam * (Ty[u] - ctr.y), z = lam * (Tz[u] - ctr.z);
            T2[u] = fma(z, z, fma(y, y, x * x));
            Tx[u] = -2.0 * x; Ty[u] = -2.0 * y; Tz[u] = -2.0 * z;
        }
    }
    // Mixed grid targets with grid_scaled: the same centred, scaled expansion
    // around the target node's centre c (the middle of its Chebyshev range).
    // Every source of a grid key passed the acceptance criterion,
    // (R_t + R_s) < theta d, so r >= (1 - theta) d while |t-c|^2 + |s-c|^2 <=
    // (theta^2 + (1 + theta)^2) d^2: for theta < 0.99 the expansion's relative
    // error stays below 4e-10 and no pair needs a check.  FP32 sums over
    // blocks of 32 sources.
    constexpr bool GS = MIXED && GRID;
    const bool gscaled = GS && A.grid_scaled;
    if constexpr (GS) {
        if (gscaled) {
            lam = KER == KY ? 1.4426950408889634 * A.kp.p0 : (SS ? sqrt(1.4426950408889634 * A.kp.inv_a2) : 1.0);
            eps2s = (float)(lam * lam * A.kp.eps2);
            ctr = make_double4(0.5 * (ttab[0] + ttab[nl - 1]), 0.5 * (ttab[nl] + ttab[2 * nl - 1]),
                               0.5 * (ttab[2 * nl] + ttab[3 * nl - 1]), 0.0);
#pragma unroll
            for (int u = 0; u < TPT; u++) {
                const double x = lam * (Tx[u] - ctr.x), y = lam * (Ty[u] - ctr.y), z = lam * (Tz[u] - ctr.z);
                T2[u] = fma(z, z, fma(y, y, x * x));
                Tx[u] = -2.0 * x; Ty[u] = -2.0 * y; Tz[u] = -2.0 * z;
            }
        }
    }

    // n staged sources against the first NU target slots
    auto consume_nu = [&](int n, auto check_tag, auto nu_tag) {
        constexpr bool CHECK = decltype(check_tag)::value;
        constexpr int NU = decltype(nu_tag)::value;
        if constexpr (GS) {
            if (gscaled) {
                constexpr int BLK = 32;
                for (int jb = 0; jb < n; jb += BLK) {
                    const int je = min(jb + BLK, n);
                    float p1[NU], p2[NU];
#pragma unroll
                    for (int u = 0; u < NU; u++) { p1[u] = 0.0f; p2[u] = 0.0f; }
                    auto body = [&](int j) {
                        const double4 s = Ls[j];
                        const float2 f = Lf[j];
#pragma unroll
                        for (int u = 0; u < NU; u++) {
                            const float r2 =
                                (float)fma(Tx[u], s.x, fma(Ty[u], s.y, fma(Tz[u], s.z, T2[u] + s.w)));
                            pair_grid_scaled<KER, SS>(r2, f.x, f.y, eps2s, p1[u], p2[u]);
                        }
                    };
                    if (je - jb == BLK) {
#pragma unroll 8
                        for (int j = 0; j < BLK; ++j) body(jb + j);
                    } else {
#pragma unroll 2
                        for (int j = jb; j < je; ++j) body(j);
                    }
#pragma unroll
                    for (int u = 0; u < NU; u++) {
                        acc[u] += (double)p1[u];
                        if constexpr (SS) accw[u] += (double)p2[u];
                    }
                }
                return;
            }
        }
        if constexpr (SCALED) {
            // CHECK: the range may hold pairs below the threshold (compared
            // in FP32; a threshold below FLT_MIN still catches r'^2 <= 0)
            constexpr int BLK = 32;
            const float thf = fmaxf((float)th, 1.17549435e-38f);
            for (int jb = 0; jb < n; jb += BLK) {
                const int je = min(jb + BLK, n);
                float p1[NU], p2[NU];
#pragma unroll
                for (int u = 0; u < NU; u++) { p1[u] = 0.0f; p2[u] = 0.0f; }
                float rmin = 3.0e38f;  // least r'^2 of the block (CHECK)
                auto body = [&](int j) {
                    const double4 s = Ls[j];
                    const float2 f = Lf[j];
#pragma unroll
                    for (int u = 0; u < NU; u++) {
                        const float r2 =
                            (float)fma(Tx[u], s.x, fma(Ty[u], s.y, fma(Tz[u], s.z, T2[u] + s.w)));
                        if constexpr (CHECK) rmin = fminf(rmin, r2);
                        pair_ss32<KER, false>(r2, f.x, f.y, eps2s, p1[u], p2[u]);
                    }
                };
                if (je - jb == BLK) {
                    // full block: constant LDS offsets from one base address
#pragma unroll 8
                    for (int j = 0; j < BLK; ++j) body(jb + j);
                } else {
#pragma unroll 2
                    for (int j = jb; j < je; ++j) body(j);
                }
                if constexpr (CHECK) {
                    if (__any(rmin < thf)) {
                        // the block again from coordinate differences (T' = -T/2 exactly)
#pragma unroll
                        for (int u = 0; u < NU; u++) { p1[u] = 0.0f; p2[u] = 0.0f; }
                        for (int j = jb; j < je; ++j) {
                            const double4 s = Ls[j];
                            const float2 f = Lf[j];
#pragma unroll
                            for (int u = 0; u < NU; u++) {
                                const double dx = fma(-0.5, Tx[u], -s.x), dy = fma(-0.5, Ty[u], -s.y),
                                             dz = fma(-0.5, Tz[u], -s.z);
                                const float r2 = (float)fma(dz, dz, fma(dy, dy, dx * dx));
                                pair_ss32<KER, KER != KR>(r2, f.x, f.y, eps2s, p1[u], p2[u]);
                            }
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < NU; u++) acc[u] += (double)p1[u] - Tq[u] * (double)p2[u];
            }
            return;
        }
#pragma unroll 2
        for (int j = 0; j < n; ++j) {
            const double4 s = Ls[j];
            const double b = SS ? Lb[j] : 0.0;
#pragma unroll
            for (int u = 0; u < NU; u++) {
                if constexpr (MIXED) {
                    // grid targets outside the scaled form (theta >= 0.99): FP64
                    // displacement and r^2, rounded once (mixed particle targets
                    // have returned from the scaled path above)
                    if constexpr (GRID) {
                        const double dx = Tx[u] - s.x, dy = Ty[u] - s.y, dz = Tz[u] - s.z;
                        const float r2 = (float)fma(dz, dz, fma(dy, dy, dx * dx));
                        pair_grid32<KER, SS>(r2, s.w, b, k32, acc[u], accw[u]);
                    }
                } else {
                    const double dx = Tx[u] - s.x, dy = Ty[u] - s.y, dz = Tz[u] - s.z;
                    if constexpr (GRID) pair_grid<KER, SS, true>(dx, dy, dz, s.w, b, A.kp, acc[u], accw[u], Ltab);
                    else pair_pot<KER, SS, CHECK, true>(dx, dy, dz, s.w, b, Tq[u], A.kp, acc[u], Ltab);
                }
            }
        }
    };
    auto consume = [&](int n, auto check_tag) {
        if (nu == TPT) consume_nu(n, check_tag, std::integral_constant<int, TPT>{});
        else if (TPT > 1 && nu == 1) consume_nu(n, check_tag, std::integral_constant<int, 1>{});
    };
    using NoCheck = std::integral_constant<bool, false>;
    // scaled: the threshold vote; otherwise the skip of coincident pairs
    using Check = std::integral_constant<bool, SCALED || KER != KR>;
    // one staged source point (thread tid)
    auto stage = [&](double x, double y, double z, double a, double b) {
        if (SCALED || gscaled) {
            x = lam * (x - ctr.x); y = lam * (y - ctr.y); z = lam * (z - ctr.z);
            Ls[tid] = make_double4(x, y, z, fma(z, z, fma(y, y, x * x)));
            Lf[tid] = make_float2((float)(lam * a), (float)(lam * b));
        } else {
            Ls[tid] = make_double4(x, y, z, a);
            if constexpr (SS) Lb[tid] = b;
        }
    };
    const bool clu_check = SCALED && can_close && ctr.w < 0.0;

    // source clusters (PC in PC mode, CC in CC mode)
    const int c0 = A.clu_off[key], c1 = A.clu_off[key + 1];
    for (int c = c0; c < c1; ++c) {
        const int id = A.clu_id[c];
        int P = A.P;
        const double *cx = A.cx, *cy = A.cy, *cz = A.cz, *cq = A.cq, *cw = A.cw;
        size_t base;
        if (A.multi) {
            const int dg = id >> DEG_SHIFT;
            P = (dg + 1) * (dg + 1) * (dg + 1);
            cx = A.dcx[dg]; cy = A.dcy[dg]; cz = A.dcz[dg]; cq = A.dcq[dg]; cw = A.dcw[dg];
            base = (size_t)(id & SLOT_MASK) * P;
        } else {
            base = (size_t)id * P;
        }
        for (int off = 0; off < P; off += TPB) {
            const int n = min(TPB, P - off);
            __syncthreads();
            if (tid < n) {
                const size_t g = base + off + tid;
                stage(cx[g], cy[g], cz[g], cq[g], SS ? cw[g] : 0.0);
            }
            __syncthreads();
            if (clu_check) consume(n, Check{});
            else consume(n, NoCheck{});
        }
    }
    // source particle ranges (PP, or CP into grids)
    const int s0 = A.seg_off[key], s1 = A.seg_off[key + 1];
    for (int s = s0; s < s1; ++s) {
        const Seg sg = A.segs[s];
        if (A.skip_sym && ((unsigned)sg.y & SEG_SYM_BIT) != 0) continue;  // the symmetric launch's
        const bool check = SCALED ? can_close && ((unsigned)sg.y & SEG_CLOSE_BIT) != 0
                                  : !GRID && ((unsigned)sg.y & SEG_COINCIDE_BIT) != 0;
        const int cnt = (int)((unsigned)sg.y & SEG_COUNT_MASK);
        for (int off = 0; off < cnt; off += TPB) {
            const int n = min(TPB, cnt - off);
            __syncthreads();
            if (tid < n) {
                const int g = sg.x + off + tid;
                const double q = A.sq[g];
                double a = q, b = 0.0;
                if constexpr (SS) { b = A.sw[g]; a = q * b; }
                stage(A.sx[g], A.sy[g], A.sz[g], a, b);
            }
            __syncthreads();
            if (check) consume(n, Check{});
            else consume(n, NoCheck{});
        }
    }
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        if (tix[u] < 0) continue;
        if constexpr (KER == KAT && !GRID) acc[u] *= A.kp.c2;  // atan_pp.c:58, atan_pc.c:58
        if constexpr (GRID) {
            const size_t o = orow + tix[u];
            A.out1[o] = acc[u];
            if constexpr (SS) A.out2[o] = accw[u];
        } else if (A.accumulate) {
            A.out1[tix[u]] += acc[u];
        } else {
            A.out1[tix[u]] = acc[u];
        }
    }
}

// ---------------------------------------------------------------------
// Grid targets as tensor grids (mixed precision, acceptance-separated
// sources: grid_scaled).  The NL^3 points of a target node's grid are
// (x_a, y_b, z_c), so r^2 = (x_a - s_x)^2 + [(y_b - s_y)^2 + (z_c - s_z)^2]
// and the bracket is shared by the NL points of the pencil (b, c).  A unit
// (thread) owns a whole pencil, so U = NL^2 units; the 256 threads form
// G = 256/U groups of U units that take every G-th staged source, and the
// groups' sums are added in group order at the end.  Per staged source a
// unit forms the bracket once (two subtractions, a multiply and an fma),
// then per point one subtraction and one fma before the kernel
// (against four FP64 instructions and a conversion per pair in the centred
// expansion of interact_tile).  Coordinates are taken relative to the node's
// centre and scaled by lam as in that expansion, then rounded to FP32; a
// difference then carries the rounding of the two centred coordinates
// (below 6e-8 of their size).  An acceptance-separated pair has r >= (1 -
// theta) d with centred coordinates below about 2 d, so r^2 keeps a
// relative error below about 5e-7 / (1 - theta), far less for most pairs.
// FP32 sums over blocks of 32 sources, FP64 totals; rows as in k_interact.
// One block per tile {key, 0, NL^3, node}; NL from the node field's degree
// (several degrees) or the launch's.
// ---------------------------------------------------------------------

// r2s: the scaled r'^2 (regularized kernel: r'^2 + e'^2)
template <int KER, bool SS>
__device__ __forceinline__ void pair_pencil(float r2s, float af, float bf, float &p1, float &p2)
{
    if constexpr (KER == KC) {
        const float ir = rsq32(r2s);
        p1 = fmaf(af, ir, p1);
        if constexpr (SS) p2 = fmaf(bf, exp2_32(-r2s) * ir, p2);
    } else if constexpr (KER == KY) {
        const float ir = rsq32(r2s);
        const float e = exp2_32(-(r2s * ir)) * ir;
        p1 = fmaf(af, e, p1);
        if constexpr (SS) p2 = fmaf(bf, e, p2);
    } else {
        p1 = fmaf(af, rsq32(r2s), p1);
    }
}

constexpr int PENCIL_RED = 1536;  // doubles per part of the group reduction: max (G-1) U NL

template <int KER, bool SS, int NL>
__device__ __forceinline__ void grid_pencil(const InteractArgs &A, const Tile tile, float4 *Ls, float *Lb,
                                            double *Lr)
{
    constexpr int P2 = NL * NL, L = NL, U = P2;
    constexpr int G = 256 / U;  // source groups
    static_assert(U * G <= 256 && G >= 1 && (G - 1) * U * L <= PENCIL_RED, "pencil layout");
    const int key = tile.x, tid = threadIdx.x;
    const double *ttab = A.multi ? A.dtab[NL - 1] + (size_t)(tile.w & SLOT_MASK) * 3 * NL
                                 : A.tnode_tab + (size_t)tile.w * 3 * NL;
    const size_t orow = A.out_row ? (size_t)A.out_row[key] : (size_t)tile.w * P2 * NL;
    const double lam = KER == KY ? 1.4426950408889634 * A.kp.p0
                                 : (SS ? sqrt(1.4426950408889634 * A.kp.inv_a2) : 1.0);
    const float eps2s = (float)(lam * lam * A.kp.eps2);
    const double cx = 0.5 * (ttab[0] + ttab[NL - 1]), cy = 0.5 * (ttab[NL] + ttab[2 * NL - 1]),
                 cz = 0.5 * (ttab[2 * NL] + ttab[3 * NL - 1]);
    const bool on = tid < U * G;
    const int grp = tid / U, unit = tid % U;
    const int pen = unit % P2, a0 = (unit / P2) * L;
    const float gy = (float)(lam * (ttab[NL + pen % NL] - cy)), gz = (float)(lam * (ttab[2 * NL + pen / NL] - cz));
    float gx[L], p1[L], p2[L];
    double acc[L], accw[L];
#pragma unroll
    for (int i = 0; i < L; i++) {
        gx[i] = (float)(lam * (ttab[min(a0 + i, NL - 1)] - cx));
        p1[i] = 0.0f; p2[i] = 0.0f; acc[i] = 0.0; accw[i] = 0.0;
    }
    auto stage = [&](double x, double y, double z, double qa, double qb) {
        Ls[tid] = make_float4((float)(lam * (x - cx)), (float)(lam * (y - cy)), (float)(lam * (z - cz)),
                              (float)(lam * qa));
        if constexpr (SS) Lb[tid] = (float)(lam * qb);
    };
    // the staged sources grp, grp + G, ... of n
    auto consume = [&](int n) {
        if (!on) return;
        for (int jb = grp; jb < n; jb += 32 * G) {
            const int je = min(jb + 32 * G, n);
            for (int j = jb; j < je; j += G) {
                const float4 s = Ls[j];
                const float bf = SS ? Lb[j] : 0.0f;
                const float dy = gy - s.y, dz = gz - s.z;
                float dyz = fmaf(dz, dz, dy * dy);
                if constexpr (KER == KR) dyz += eps2s;
#pragma unroll
                for (int i = 0; i < L; i++) {
                    const float dx = gx[i] - s.x;
                    pair_pencil<KER, SS>(fmaf(dx, dx, dyz), s.w, bf, p1[i], p2[i]);
                }
            }
#pragma unroll
            for (int i = 0; i < L; i++) {
                acc[i] += (double)p1[i];
                p1[i] = 0.0f;
                if constexpr (SS) { accw[i] += (double)p2[i]; p2[i] = 0.0f; }
            }
        }
    };
    // source clusters (CC)
    for (int c = A.clu_off[key]; c < A.clu_off[key + 1]; ++c) {
        const int id = A.clu_id[c];
        int P = A.P;
        const double *sx = A.cx, *sy = A.cy, *sz = A.cz, *sq = A.cq, *sw = A.cw;
        size_t base;
        if (A.multi) {
            const int dg = id >> DEG_SHIFT;
            P = (dg + 1) * (dg + 1) * (dg + 1);
            sx = A.dcx[dg]; sy = A.dcy[dg]; sz = A.dcz[dg]; sq = A.dcq[dg]; sw = A.dcw[dg];
            base = (size_t)(id & SLOT_MASK) * P;
        } else {
            base = (size_t)id * P;
        }
        for (int off = 0; off < P; off += 256) {
            const int n = min(256, P - off);
            __syncthreads();
            if (tid < n) {
                const size_t g = base + off + tid;
                stage(sx[g], sy[g], sz[g], sq[g], SS ? sw[g] : 0.0);
            }
            __syncthreads();
            consume(n);
        }
    }
    // source particle ranges (CP)
    for (int s = A.seg_off[key]; s < A.seg_off[key + 1]; ++s) {
        const Seg sg = A.segs[s];
        const int cnt = (int)((unsigned)sg.y & SEG_COUNT_MASK);
        for (int off = 0; off < cnt; off += 256) {
            const int n = min(256, cnt - off);
            __syncthreads();
            if (tid < n) {
                const int g = sg.x + off + tid;
                const double q = A.sq[g];
                double qa = q, qb = 0.0;
                if constexpr (SS) { qb = A.sw[g]; qa = q * qb; }
                stage(A.sx[g], A.sy[g], A.sz[g], qa, qb);
            }
            __syncthreads();
            consume(n);
        }
    }
    if constexpr (G > 1) {
        // the groups' sums, added in group order
        if (on && grp > 0)
#pragma unroll
            for (int i = 0; i < L; i++) {
                Lr[((grp - 1) * U + unit) * L + i] = acc[i];
                if constexpr (SS) Lr[PENCIL_RED + ((grp - 1) * U + unit) * L + i] = accw[i];
            }
        __syncthreads();
        if (on && grp == 0)
            for (int g = 1; g < G; g++)
#pragma unroll
                for (int i = 0; i < L; i++) {
                    acc[i] += Lr[((g - 1) * U + unit) * L + i];
                    if constexpr (SS) accw[i] += Lr[PENCIL_RED + ((g - 1) * U + unit) * L + i];
                }
    }
    if (!on || grp != 0) return;
#pragma unroll
    for (int i = 0; i < L; i++) {
        if (a0 + i >= NL) continue;
        const size_t o = orow + a0 + i + (size_t)NL * pen;
        A.out1[o] = acc[i];
        if constexpr (SS) A.out2[o] = accw[i];
    }
}

template <int KER, bool SS>
__global__ __launch_bounds__(256) void k_grid_pencil(InteractArgs A)
{
    __shared__ float4 Ls[256];          // S' (centred, scaled) and the scaled a charge
    __shared__ float Lb[SS ? 256 : 1];  // the scaled b charge (subtraction)
    __shared__ double Lr[(SS ? 2 : 1) * PENCIL_RED];
    const Tile tile = A.tiles[blockIdx.x];
    if (tile.z == 0) return;  // padding (block-uniform)
    const int nl = A.multi ? (tile.w >> DEG_SHIFT) + 1 : A.nl;
    switch (nl) {
        case 2: grid_pencil<KER, SS, 2>(A, tile, Ls, Lb, Lr); break;
        case 3: grid_pencil<KER, SS, 3>(A, tile, Ls, Lb, Lr); break;
        case 4: grid_pencil<KER, SS, 4>(A, tile, Ls, Lb, Lr); break;
        case 5: grid_pencil<KER, SS, 5>(A, tile, Ls, Lb, Lr); break;
        case 6: grid_pencil<KER, SS, 6>(A, tile, Ls, Lb, Lr); break;
        case 7: grid_pencil<KER, SS, 7>(A, tile, Ls, Lb, Lr); break;
        case 8: grid_pencil<KER, SS, 8>(A, tile, Ls, Lb, Lr); break;
        case 9: grid_pencil<KER, SS, 9>(A, tile, Ls, Lb, Lr); break;
        case 10: grid_pencil<KER, SS, 10>(A, tile, Ls, Lb, Lr); break;
        case 11: grid_pencil<KER, SS, 11>(A, tile, Ls, Lb, Lr); break;
        case 12: grid_pencil<KER, SS, 12>(A, tile, Ls, Lb, Lr); break;
        case 13: grid_pencil<KER, SS, 13>(A, tile, Ls, Lb, Lr); break;
        default: break;
    }
}

// ---------------------------------------------------------------------
// FP64 executions on target grids as tensor grids: the pencils and source
// groups of grid_pencil, with the reference's displacements (grid point
// minus source, in FP64) and pair expressions (grid_term; coulomb_cp.c:46-55,
// coulomb_ss_cp.c:54-55, yukawa_cp.c:52, yukawa_ss_cp.c:55-56,
// regularized-coulomb_cp.c:52 and the CC forms).  dy^2 + dz^2 (+ e^2 for the
// regularized kernel) is formed once per pencil and source, so a pair costs
// one subtraction and one fma before the kernel instead of three
// subtractions, a product and two fma.  r^2 is summed as dx^2 + (dz^2 +
// dy^2) where the reference sums dx^2 + dy^2 + dz^2: the two differ by
// rounding (a few ulp of r^2).  FP64 sums per thread; the groups' sums are
// added in group order.
// ---------------------------------------------------------------------
template <int KER, bool SS, int NL>
__device__ __forceinline__ void grid_pencil64(const InteractArgs &A, const Tile tile, double4 *Ls, double *Lb,
                                              double *Lr, const double *Ltab)
{
    constexpr int U = NL * NL, L = NL;
    constexpr int G = 256 / U;  // source groups
    static_assert(U * G <= 256 && G >= 1 && (G - 1) * U * L <= PENCIL_RED, "pencil layout");
    const int key = tile.x, tid = threadIdx.x;
    const double *ttab = A.multi ? A.dtab[NL - 1] + (size_t)(tile.w & SLOT_MASK) * 3 * NL
                                 : A.tnode_tab + (size_t)tile.w * 3 * NL;
    const size_t orow = A.out_row ? (size_t)A.out_row[key] : (size_t)tile.w * U * NL;
    const bool on = tid < U * G;
    const int grp = tid / U, pen = tid % U;
    const double gy = ttab[NL + pen % NL], gz = ttab[2 * NL + pen / NL];
    const double e2 = KER == KR ? A.kp.eps2 : 0.0;
    double gx[L], acc[L], accw[L];
#pragma unroll
    for (int i = 0; i < L; i++) {
        gx[i] = ttab[i];
        acc[i] = 0.0; accw[i] = 0.0;
    }
    // the staged sources grp, grp + G, ... of n
    auto consume = [&](int n) {
        if (!on) return;
        for (int j = grp; j < n; j += G) {
            const double4 s = Ls[j];
            const double b = SS ? Lb[j] : 0.0;
            const double dy = gy - s.y, dz = gz - s.z;
            double dyz = fma(dz, dz, dy * dy);
            if constexpr (KER == KR) dyz += e2;
#pragma unroll
            for (int i = 0; i < L; i++) {
                const double dx = gx[i] - s.x;
                grid_term<KER, SS, true, KER == KR>(fma(dx, dx, dyz), s.w, b, A.kp, acc[i], accw[i], Ltab);
            }
        }
    };
    auto stage = [&](double x, double y, double z, double qa, double qb) {
        Ls[tid] = make_double4(x, y, z, qa);
        if constexpr (SS) Lb[tid] = qb;
    };
    // source clusters (CC)
    for (int c = A.clu_off[key]; c < A.clu_off[key + 1]; ++c) {
        const int id = A.clu_id[c];
        int P = A.P;
        const double *sx = A.cx, *sy = A.cy, *sz = A.cz, *sq = A.cq, *sw = A.cw;
        size_t base;
        if (A.multi) {
            const int dg = id >> DEG_SHIFT;
            P = (dg + 1) * (dg + 1) * (dg + 1);
            sx = A.dcx[dg]; sy = A.dcy[dg]; sz = A.dcz[dg]; sq = A.dcq[dg]; sw = A.dcw[dg];
            base = (size_t)(id & SLOT_MASK) * P;
        } else {
            base = (size_t)id * P;
        }
        for (int off = 0; off < P; off += 256) {
            const int n = min(256, P - off);
            __syncthreads();
            if (tid < n) {
                const size_t g = base + off + tid;
                stage(sx[g], sy[g], sz[g], sq[g], SS ? sw[g] : 0.0);
            }
            __syncthreads();
            consume(n);
        }
    }
    // source particle ranges (CP)
    for (int s = A.seg_off[key]; s < A.seg_off[key + 1]; ++s) {
        const Seg sg = A.segs[s];
        const int cnt = (int)((unsigned)sg.y & SEG_COUNT_MASK);
        for (int off = 0; off < cnt; off += 256) {
            const int n = min(256, cnt - off);
            __syncthreads();
            if (tid < n) {
                const int g = sg.x + off + tid;
                const double q = A.sq[g];
                double qa = q, qb = 0.0;
                if constexpr (SS) { qb = A.sw[g]; qa = q * qb; }
                stage(A.sx[g], A.sy[g], A.sz[g], qa, qb);
            }
            __syncthreads();
            consume(n);
        }
    }
    if constexpr (G > 1) {
        // the groups' sums, added in group order
        if (on && grp > 0)
#pragma unroll
            for (int i = 0; i < L; i++) {
                Lr[((grp - 1) * U + pen) * L + i] = acc[i];
                if constexpr (SS) Lr[PENCIL_RED + ((grp - 1) * U + pen) * L + i] = accw[i];
            }
        __syncthreads();
        if (on && grp == 0)
            for (int g = 1; g < G; g++)
#pragma unroll
                for (int i = 0; i < L; i++) {
                    acc[i] += Lr[((g - 1) * U + pen) * L + i];
                    if constexpr (SS) accw[i] += Lr[PENCIL_RED + ((g - 1) * U + pen) * L + i];
                }
    }
    if (!on || grp != 0) return;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const size_t o = orow + i + (size_t)NL * pen;
        A.out1[o] = acc[i];
        if constexpr (SS) A.out2[o] = accw[i];
    }
}

template <int KER, bool SS>
__global__ __launch_bounds__(256) void k_grid_pencil64(InteractArgs A)
{
    __shared__ double4 Ls[256];          // source point and its a charge
    __shared__ double Lb[SS ? 256 : 1];  // its b charge (subtraction)
    __shared__ double Lr[(SS ? 2 : 1) * PENCIL_RED];
    __shared__ double Ltab[EXP_TAB];     // 2^(j/64) for exp_nonpos
    if constexpr (SS || KER == KY) exp_table(Ltab, threadIdx.x);  // read after the first staging barrier
    const Tile tile = A.tiles[blockIdx.x];
    if (tile.z == 0) return;  // padding (block-uniform)
    const int nl = A.multi ? (tile.w >> DEG_SHIFT) + 1 : A.nl;
    switch (nl) {
        case 2: grid_pencil64<KER, SS, 2>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 3: grid_pencil64<KER, SS, 3>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 4: grid_pencil64<KER, SS, 4>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 5: grid_pencil64<KER, SS, 5>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 6: grid_pencil64<KER, SS, 6>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 7: grid_pencil64<KER, SS, 7>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 8: grid_pencil64<KER, SS, 8>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 9: grid_pencil64<KER, SS, 9>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 10: grid_pencil64<KER, SS, 10>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 11: grid_pencil64<KER, SS, 11>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 12: grid_pencil64<KER, SS, 12>(A, tile, Ls, Lb, Lr, Ltab); break;
        case 13: grid_pencil64<KER, SS, 13>(A, tile, Ls, Lb, Lr, Ltab); break;
        default: break;
    }
}

// Interaction kernel.  Tiles are laid out as 8 interleaved queues: tile
// 8*j + x is the j-th tile of queue x, and each queue's tiles read sources
// that its neighbours in the queue also read.  Workgroups are dispatched
// round-robin over the 8 XCDs, so block b runs on XCD b % 8.
//   one-shot  : one block per tile slot (padding slots have count 0);
//   persistent: a grid sized to the resident capacity; block b pops tiles
//               from queue b % 8 through an atomic cursor (vector memory
//               atomics), then steals from the other queues, and exits when
//               every queue is empty -- every block reaches that exit.
template <int KER, bool SS, bool GRID, bool MIXED, int TPT>
__global__ __launch_bounds__(TPB) void k_interact(InteractArgs A)
{
    __shared__ double4 Ls[TPB];
    __shared__ double Lb[SS ? TPB : 1];
    __shared__ float2 Lf[MIXED && (SS || GRID) ? TPB : 1];
    // 2^(j/64) for the FP64 exponentials (read after the first staging barrier)
    __shared__ double Ltab[EXP_TAB];
    if constexpr (!MIXED && (SS || KER == KY)) exp_table(Ltab, threadIdx.x);
    if (!A.persistent) {
        const Tile tile = A.tiles[blockIdx.x];
        if (tile.z == 0) return;  // padding tile (uniform across the block)
        interact_tile<KER, SS, GRID, MIXED, TPT>(A, tile, Ls, Lb, Lf, Ltab);
        return;
    }
    __shared__ int Lt;
    const int home = blockIdx.x % NQ;
    for (;;) {
        if (threadIdx.x == 0) {
            int t = -1;
            for (int k = 0; k < NQ && t < 0; k++) {
                const int q = (home + k) % NQ;
                const int len = A.qlen[q];
                if (__hip_atomic_load(&A.cursor[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len)
                    continue;
                const int j = atomicAdd(&A.cursor[q], 1);
                if (j < len) t = j * NQ + q;
            }
            Lt = t;
        }
        __syncthreads();
        const int t = Lt;
        __syncthreads();
        if (t < 0) break;
        interact_tile<KER, SS, GRID, MIXED, TPT>(A, A.tiles[t], Ls, Lb, Lf, Ltab);
    }
}

// ---------------------------------------------------------------------
// Symmetric near field.  One block per item (leaves A, B).  Coordinates are
// taken relative to the item's centre c and scaled by lam as in pair_ss32
// (charges by lam), r'^2 = |T'|^2 + |S'|^2 - 2 T'.S' in FP64 as in the
// scaled path of interact_tile.  Lanes form groups of G: a pass holds TPT
// targets of A per lane, lane l of wave w owning targets
// w*G*TPT + (l%G) + G*k, k < TPT (the 64/G lanes with the same l%G hold the
// same targets), and a staged chunk of up to 64*NW sources of B is split over
// the groups: at step s lane l takes source (64/G)*s + l/G against its TPT
// targets.  The targets' sums (a terms and b terms apart, FP32, flushed to
// FP64 every 32 steps) stay in registers over the whole of B.  A source's
// sum over a lane's targets (subtraction: the a terms minus the source's
// charge times the b terms, combined in FP32 before any lane exchange, so
// one scan per step serves both parts) is added over its group (DPP row_shr
// scan: the group's last lane ends with the sum), over the NW waves in LDS
// in FP64, and into the source's column slot (stored at the first pass over
// A, added after: the block owns its slots, so the order is fixed).  In
// items flagged close, a wave whose chunk held a pair below the threshold
// evaluates the chunk again from coordinate differences (coincident pairs
// add exactly zero, the reference's skip).  One-sided items run the same
// loop without the sources' sums.
// ---------------------------------------------------------------------
constexpr int SYM_G = 8, SYM_NW = 4;
// targets per lane: 6 for Coulomb and Yukawa (the per-step scan and loop
// cost spread over more pairs; up to 155 VGPRs, three waves per SIMD: 1-3%
// faster than 4 targets at four waves), 4 for the regularized kernel (whose
// one-sided items, most of its pairs at 4M Plummer, ran 2.5% slower with 6)
constexpr int sym_tpt(int ker, bool) { return ker == KR ? 4 : 6; }
constexpr int sym_wpe(int ker, bool ss) { return sym_tpt(ker, ss) == 6 ? 3 : 4; }

// Sums over groups of G consecutive lanes (G = 2, 4, 8 or 16) by a
// row_shr scan inside each DPP row of 16 lanes: after the shifts 1, 2, ..,
// G/2 the last lane of every group holds the group's sum.
template <int G>
__device__ __forceinline__ float group_sum(float v)
{
    // bound_ctrl: a lane whose source lane lies outside its row reads 0
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, true));
    if constexpr (G > 2)
        v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xf, 0xf, true));
    if constexpr (G > 4)
        v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xf, 0xf, true));
    if constexpr (G > 8)
        v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xf, 0xf, true));
    return v;
}

// One pair's terms for both directions, from r'^2 (scaled as in pair_ss32
// for subtraction; skipping: lam = kappa log2(e) for Yukawa, 1 otherwise):
// the a-part factor fa and the b-part factor fb (subtraction only).  SKIP:
// a coincident pair (r^2 = 0) gives zero, the reference's skip (the
// regularized kernel has none).  EPS_IN: r2 of the regularized kernel
// already holds e'^2.
template <int KER, bool SS, bool SKIP, bool EPS_IN = false>
__device__ __forceinline__ void sym_pair(float r2, float eps2s, float &fa, float &fb)
{
    if constexpr (KER == KR) {
        fa = rsq32(EPS_IN ? r2 : r2 + eps2s);
        if constexpr (SS) fb = exp2_32(-r2) * fa;
        return;
    }
    float ir = rsq32(r2);
    if constexpr (KER == KC) {
        if constexpr (SS) fb = exp2_32(-r2) * ir;
        fa = ir;
    } else {
        fa = exp2_32(-(r2 * ir)) * ir;
        fb = fa;
    }
    if constexpr (SKIP) {
        fa = r2 > 0.0f ? fa : 0.0f;
        if constexpr (SS) fb = r2 > 0.0f ? fb : 0.0f;
    }
}

// a source's a and b (scaled) and, subtraction, its charge as a target
template <bool SS> struct SymSrc { float a, b, q, pad; };
template <> struct SymSrc<false> { float a, b; };

template <int KER, bool SS, int TPT, int G, int NW, int WPE>
__global__ __launch_bounds__(64 * NW, WPE) void k_near_sym(SymArgs A)  // at least WPE waves per SIMD
{
    constexpr int NT = 64 * NW;  // threads of the block; sources per staged chunk
    __shared__ double4 Ls[NT];     // S' and |S'|^2
    __shared__ SymSrc<SS> Lf[NT];
    __shared__ float Lc[NW][NT];   // per wave: the sources' sums
    const SymItem it = A.items[blockIdx.x];
    const double4 ctr = A.ctr[blockIdx.x];
    constexpr int NG = 64 / G;  // source groups of a wave
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, rg = lane % G, cg = lane / G;
    const double lam = KER == KY ? 1.4426950408889634 * A.kp.p0
                                 : (SS ? sqrt(1.4426950408889634 * A.kp.inv_a2) : 1.0);
    const float eps2s = (float)(lam * lam * A.kp.eps2);
    // the regularized kernel's sensitivity is to r^2 + eps^2, which its
    // expanded r'^2 holds (e'^2 folded into |T'|^2: no FP32 add per pair)
    constexpr bool EPS_IN = KER == KR;
    const double th = lam * lam * (fabs(ctr.w) - (KER == KR ? A.kp.eps2 : 0.0));
    const bool check = ctr.w < 0.0 && th > 0.0;
    const float thf = fmaxf((float)(EPS_IN ? lam * lam * fabs(ctr.w) : th), 1.17549435e-38f);
    constexpr int TPASS = NW * G * TPT;
    // a and b of a particle as a source: subtraction (lam q w, lam w),
    // skipping (lam q, 0)
    auto charges = [&](int g) {
        const double q = A.q[g];
        if constexpr (SS) {
            const double w = A.w[g];
            return SymSrc<true>{(float)(lam * q * w), (float)(lam * w), (float)q, 0.0f};
        } else {
            return SymSrc<false>{(float)(lam * q), 0.0f};
        }
    };

    // TWO: the sources' sums as well (block-uniform)
    auto body = [&](auto two_tag) {
    constexpr bool TWO = decltype(two_tag)::value;
    for (int p0 = 0; p0 < it.a_cnt; p0 += TPASS) {
        double Tx[TPT], Ty[TPT], Tz[TPT], T2[TPT], Tq[TPT], acc[TPT];
        float ai[TPT], bi[TPT], p1[TPT], p2[TPT];
#pragma unroll
        for (int k = 0; k < TPT; k++) {
            const int li = p0 + wv * G * TPT + rg + G * k;
            const bool v = li < it.a_cnt;
            const int g = it.a_beg + (v ? li : 0);
            const double x = lam * (A.tx[g] - ctr.x), y = lam * (A.ty[g] - ctr.y), z = lam * (A.tz[g] - ctr.z);
            T2[k] = fma(z, z, fma(y, y, x * x)) + (EPS_IN ? lam * lam * A.kp.eps2 : 0.0);
            Tx[k] = -2.0 * x; Ty[k] = -2.0 * y; Tz[k] = -2.0 * z;
            Tq[k] = SS ? A.tq[g] : 0.0;
            if constexpr (TWO) {
                // the target as a source (two-sided items: targets are the
                // sources); a target outside A gives nothing to the sources
                const SymSrc<SS> c = charges(g);
                ai[k] = v ? c.a : 0.0f;
                bi[k] = v ? c.b : 0.0f;
            } else {
                ai[k] = 0.0f; bi[k] = 0.0f;
            }
            acc[k] = 0.0;
            p1[k] = 0.0f; p2[k] = 0.0f;
        }
        for (int off = 0; off < it.b_cnt; off += NT) {
            const int n = min(NT, it.b_cnt - off);
            __syncthreads();  // the previous chunk's LDS has been read
            if (tid < n) {
                const int g = it.b_beg + off + tid;
                const double x = lam * (A.x[g] - ctr.x), y = lam * (A.y[g] - ctr.y), z = lam * (A.z[g] - ctr.z);
                Lf[tid] = charges(g);
                Ls[tid] = make_double4(x, y, z, fma(z, z, fma(y, y, x * x)));
            } else {
                // dummies: zero charge, |S'|^2 = 1e30 (every term exactly zero)
                Ls[tid] = make_double4(0.0, 0.0, 0.0, 1.0e30);
                Lf[tid] = SymSrc<SS>{};
            }
            __syncthreads();
            const int steps = (n + NG - 1) / NG;
            double acc0[TPT];
#pragma unroll
            for (int k = 0; k < TPT; k++) acc0[k] = acc[k];
            float rmin = 3.0e38f;
            // EXACT: r^2 from coordinate differences; TRACK: the least r^2
            auto run = [&](auto exact_tag, auto track_tag) {
                constexpr bool EXACT = decltype(exact_tag)::value;
                constexpr bool TRACK = decltype(track_tag)::value;
                for (int sb = 0; sb < steps; sb += 32) {
                const int se = min(steps, sb + 32);
                for (int st = sb; st < se; st++) {
                    const int j = NG * st + cg;
                    const double4 S = Ls[j];
                    const SymSrc<SS> f = Lf[j];
                    [[maybe_unused]] float c1 = 0.0f, c2 = 0.0f;
#pragma unroll
                    for (int k = 0; k < TPT; k++) {
                        float r2;
                        if constexpr (!EXACT) {
                            r2 = (float)fma(Tx[k], S.x, fma(Ty[k], S.y, fma(Tz[k], S.z, T2[k] + S.w)));
                            if constexpr (TRACK) rmin = fminf(rmin, r2);
                        } else {
                            const double dx = fma(-0.5, Tx[k], -S.x), dy = fma(-0.5, Ty[k], -S.y),
                                         dz = fma(-0.5, Tz[k], -S.z);
                            r2 = (float)fma(dz, dz, fma(dy, dy, dx * dx));
                        }
                        float fa, fb = 0.0f;
                        sym_pair<KER, SS, EXACT && KER != KR, EPS_IN && !EXACT>(r2, eps2s, fa, fb);
                        p1[k] = fmaf(f.a, fa, p1[k]);
                        if constexpr (TWO) c1 = fmaf(ai[k], fa, c1);
                        if constexpr (SS) {
                            p2[k] = fmaf(f.b, fb, p2[k]);
                            if constexpr (TWO) c2 = fmaf(bi[k], fb, c2);
                        }
                    }
                    if constexpr (TWO) {
                        // the source's sum over the lane's targets (subtraction:
                        // a part minus its charge times the b part), then over the group
                        float c = c1;
                        if constexpr (SS) c = fmaf(-f.q, c2, c1);
                        c = group_sum<G>(c);
                        if (rg == G - 1) Lc[wv][j] = c;
                    }
                }
                // the FP32 sums of (at most) 32 steps into FP64
#pragma unroll
                for (int k = 0; k < TPT; k++) {
                    acc[k] += SS ? (double)p1[k] - Tq[k] * (double)p2[k] : (double)p1[k];
                    p1[k] = 0.0f; p2[k] = 0.0f;
                }
                }
            };
            using F = std::integral_constant<bool, false>;
            using T = std::integral_constant<bool, true>;
            if (!check) {
                run(F{}, F{});
            } else {
                run(F{}, T{});
                if (__any(rmin < thf)) {
#pragma unroll
                    for (int k = 0; k < TPT; k++) { acc[k] = acc0[k]; p1[k] = 0.0f; p2[k] = 0.0f; }
                    run(T{}, F{});
                }
            }
            if constexpr (TWO) {
                __syncthreads();
                if (tid < n) {
                    double v = 0.0;
#pragma unroll
                    for (int w = 0; w < NW; w++) v += (double)Lc[w][tid];
                    double *slot = A.slots + it.cslot + off + tid;
                    *slot = p0 == 0 ? v : *slot + v;
                }
            }
        }
        // a target's sum over the lanes that hold it
#pragma unroll
        for (int k = 0; k < TPT; k++) {
            double r = acc[k];
#pragma unroll
            for (int o = G; o < 64; o *= 2) r += __shfl_xor(r, o);
            const int li = p0 + wv * G * TPT + rg + G * k;
            if (cg == 0 && li < it.a_cnt) A.slots[it.rslot + li] = r;
        }
    }
    };
    if (it.cslot >= 0) body(std::true_type{});
    else body(std::false_type{});
}

// ---------------------------------------------------------------------
// Symmetric near field in FP64 (plain executions).  The item layout, lane
// groups and slots are those of k_near_sym; the arithmetic is that of
// pair_pot, evaluated once per unordered pair for both particles: the
// displacement of the reverse pair is the negated one, so r^2, the rsq64
// estimate and the exponential are the same numbers in both directions,
// and each direction's term is formed as pair_pot forms it with its own
// target charge.  Only the order of the sums differs from the
// one-directional launch.  Sums stay FP64 throughout: a target's over all of
// B in registers, a source's over its group's lanes by a DPP row_shr scan
// of the two 32-bit halves, then over the waves in LDS.  CHECK (items whose
// boxes touch) applies the reference's skip of coincident pairs.  One-sided
// items (a leaf and a source range that does not list it, or itself) form
// the targets' sums only.
// ---------------------------------------------------------------------
constexpr int SYM64_TPT = 4, SYM64_G = 8, SYM64_NW = 4;

template <int G>
__device__ __forceinline__ double group_sum64(double v)
{
    auto shr = [](double x, auto ctl_tag) {
        constexpr int CTL = decltype(ctl_tag)::value;
        const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTL, 0xf, 0xf, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTL, 0xf, 0xf, true);
        return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    };
    v += shr(v, std::integral_constant<int, 0x111>{});
    if constexpr (G > 2) v += shr(v, std::integral_constant<int, 0x112>{});
    if constexpr (G > 4) v += shr(v, std::integral_constant<int, 0x114>{});
    if constexpr (G > 8) v += shr(v, std::integral_constant<int, 0x118>{});
    return v;
}

template <int KER, bool SS, int TPT, int G, int NW>
__global__ __launch_bounds__(64 * NW) void k_near_sym64(SymArgs A)
{
    constexpr int NT = 64 * NW;
    __shared__ double4 Ls[NT];          // x, y, z and the source's a (q w or q)
    __shared__ double2 Lw[SS ? NT : 1]; // the source's w and its charge as a target
    __shared__ double Lc[NW][NT];       // per wave: the sources' sums
    __shared__ double Ltab[EXP_TAB];     // 2^(j/64) for exp_nonpos
    const SymItem it = A.items[blockIdx.x];
    const bool check = A.ctr[blockIdx.x].w < 0.0;
    constexpr int NG = 64 / G;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, rg = lane % G, cg = lane / G;
    constexpr int TPASS = NW * G * TPT;
    const KParams kp = A.kp;
    // read after the first chunk's barriers
    if constexpr (SS || KER == KY) exp_table(Ltab, tid);

    // TWO: the sources' sums as well (block-uniform)
    auto body = [&](auto two_tag) {
    constexpr bool TWO = decltype(two_tag)::value;
    for (int p0 = 0; p0 < it.a_cnt; p0 += TPASS) {
        // targets: coordinates, -q (subtraction), a = q w (q), w; acc
        double Tx[TPT], Ty[TPT], Tz[TPT], Tnq[TPT], Ta[TPT], Tw[TPT], acc[TPT];
#pragma unroll
        for (int k = 0; k < TPT; k++) {
            const int li = p0 + wv * G * TPT + rg + G * k;
            const bool v = li < it.a_cnt;
            const int g = it.a_beg + (v ? li : 0);
            Tx[k] = A.tx[g]; Ty[k] = A.ty[g]; Tz[k] = A.tz[g];
            Tnq[k] = SS ? -A.tq[g] : 0.0;
            if constexpr (TWO) {
                // the target as a source (two-sided items: targets are the
                // sources); a target outside A gives nothing to the sources
                const double q = A.q[g];
                const double w = SS ? A.w[g] : 1.0;
                Ta[k] = v ? (SS ? q * w : q) : 0.0;
                Tw[k] = v ? w : 0.0;
            } else {
                Ta[k] = 0.0; Tw[k] = 0.0;
            }
            acc[k] = 0.0;
        }
        for (int off = 0; off < it.b_cnt; off += NT) {
            const int n = min(NT, it.b_cnt - off);
            __syncthreads();  // the previous chunk's LDS has been read
            if (tid < n) {
                const int g = it.b_beg + off + tid;
                const double q = A.q[g];
                if constexpr (SS) {
                    const double w = A.w[g];
                    Ls[tid] = make_double4(A.x[g], A.y[g], A.z[g], q * w);
                    Lw[tid] = make_double2(w, q);
                } else {
                    Ls[tid] = make_double4(A.x[g], A.y[g], A.z[g], q);
                }
            } else {
                // dummies: zero charge and weight, far away (their column
                // sums are never stored)
                Ls[tid] = make_double4(1.0e100, 1.0e100, 1.0e100, 0.0);
                if constexpr (SS) Lw[tid] = make_double2(0.0, 0.0);
            }
            __syncthreads();
            const int steps = (n + NG - 1) / NG;
            auto run = [&](auto check_tag) {
                constexpr bool CHECK = decltype(check_tag)::value;
                for (int st = 0; st < steps; st++) {
                    const int j = NG * st + cg;
                    const double4 S = Ls[j];
                    double sw = 1.0, sq = 0.0;
                    if constexpr (SS) { const double2 u = Lw[j]; sw = u.x; sq = u.y; }
                    [[maybe_unused]] double c = 0.0;
#pragma unroll
                    for (int k = 0; k < TPT; k++) {
                        const double dx = Tx[k] - S.x, dy = Ty[k] - S.y, dz = Tz[k] - S.z;
                        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                        if constexpr (KER == KC) {
                            double ir = rsq64(r2);
                            if constexpr (CHECK) ir = r2 > (SS ? 0.0 : DBL_MIN) ? ir : 0.0;
                            if constexpr (!SS) {
                                acc[k] = fma(S.w, ir, acc[k]);
                                if constexpr (TWO) c = fma(Ta[k], ir, c);
                            } else {
                                const double e = exp_nonpos(-r2 * kp.inv_a2, Ltab);
                                acc[k] = fma(fma(Tnq[k] * sw, e, S.w), ir, acc[k]);
                                if constexpr (TWO) c = fma(fma(-sq * Tw[k], e, Ta[k]), ir, c);
                            }
                        } else if constexpr (KER == KY) {
                            const double ir = rsq64(r2);
                            double e = exp_nonpos(-kp.p0 * (r2 * ir), Ltab) * ir;
                            if constexpr (CHECK) e = r2 > 0.0 ? e : 0.0;
                            if constexpr (!SS) {
                                acc[k] = fma(S.w, e, acc[k]);
                                if constexpr (TWO) c = fma(Ta[k], e, c);
                            } else {
                                acc[k] = fma(fma(Tnq[k], sw, S.w), e, acc[k]);
                                if constexpr (TWO) c = fma(fma(-sq, Tw[k], Ta[k]), e, c);
                            }
                        } else {  // regularized Coulomb, skipping
                            const double ir = rsq64(r2 + kp.eps2);
                            acc[k] = fma(S.w, ir, acc[k]);
                            if constexpr (TWO) c = fma(Ta[k], ir, c);
                        }
                    }
                    if constexpr (TWO) {
                        c = group_sum64<G>(c);
                        if (rg == G - 1) Lc[wv][j] = c;
                    }
                }
            };
            if (check && KER != KR) run(std::integral_constant<bool, true>{});
            else run(std::integral_constant<bool, false>{});
            if constexpr (TWO) {
                __syncthreads();
                if (tid < n) {
                    double s = 0.0;
#pragma unroll
                    for (int w = 0; w < NW; w++) s += Lc[w][tid];
                    double *slot = A.slots + it.cslot + off + tid;
                    *slot = p0 == 0 ? s : *slot + s;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < TPT; k++) {
            double r = acc[k];
#pragma unroll
            for (int o = G; o < 64; o *= 2) r += __shfl_xor(r, o);
            const int li = p0 + wv * G * TPT + rg + G * k;
            if (cg == 0 && li < it.a_cnt) A.slots[it.rslot + li] = r;
        }
    }
    };
    if (it.cslot >= 0) body(std::true_type{});
    else body(std::false_type{});
}

__global__ __launch_bounds__(TPB_SYM_REDUCE) void k_sym_reduce(SymReduceArgs A)
{
    const Tile t = A.tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const int i = t.y + threadIdx.x;
    double s = A.pot[i];
    for (int k = A.off[t.x]; k < A.off[t.x + 1]; k++) s += A.slots[A.base[k] + i];
    A.pot[i] = s;
}

// ---------------------------------------------------------------------
// Near field in mixed precision with r^2 on the matrix cores.  Each wave
// owns 16*NB targets of the tile (NB row blocks of 16).  For a chunk of 16
// staged sources, v_mfma_f64_16x16x4f64 forms D = A B + C with
//   A row i  = [-2(t_i - c), 1]      (lane l supplies A[l%16][l/16])
//   B col j  = [s_j - c, |s_j - c|^2] (lane l supplies B[l/16][l%16])
//   C[i][j]  = |t_i - c|^2
// so D[i][j] = |t_i - s_j|^2 in FP64, c the key's centre; lane l holds
// D[l/16 + 4v][l%16], v = 0..3.  The FP64 rounding error of D is below
// 6e-15 (|t-c|^2 + |s-c|^2), so pairs with D below the key's threshold
// (|key_ctr.w|, 2e-7 R^2 for R the key's extent around c; less eps^2 for the
// regularized kernel, whose sensitivity is to r^2 + eps^2) take the
// coordinate differences instead, and a coincident pair (r^2 = 0) is
// skipped as the reference skips it.  The VALU then only
// rounds r^2 to FP32, evaluates the pair kernel and accumulates in FP64.
// Each target's sum is spread over the 16 lanes of its row group and added
// with a butterfly at the end of the tile.
// ---------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

template <int KER, bool SS, int NB>
__device__ __forceinline__ void near_mfma_tile(const InteractArgs &A, const Tile tile, double *LB,
                                               double *LA, float2 *LC, double *LT)
{
    constexpr int WT = 16 * NB;  // targets per wave
    constexpr int TT = 4 * WT;   // targets per tile
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int kk = lane >> 4, ii = lane & 15;
    const int key = tile.x;
    const double4 ctr = A.key_ctr[key];
    const KP32 k32{(float)(-1.4426950408889634 * A.kp.inv_a2), (float)(-1.4426950408889634 * A.kp.p0),
                   (float)A.kp.eps2};
    const double th = fabs(ctr.w) - (KER == KR ? A.kp.eps2 : 0.0);
    const bool can_close = th > 0.0;
    // centred targets of the tile (LT[d*TT + i])
    for (int i = tid; i < TT; i += TPB) {
        const int g = tile.y + (i < tile.z ? i : 0);
        LT[i] = A.tx[g] - ctr.x;
        LT[TT + i] = A.ty[g] - ctr.y;
        LT[2 * TT + i] = A.tz[g] - ctr.z;
    }
    __syncthreads();
    const int tb = wv * WT;  // first target of this wave in the tile
    const int nbv = __builtin_amdgcn_readfirstlane(max(0, min(NB, (tile.z - tb + 15) / 16)));
    double Aop[NB];
    v4d Cop[NB];
    float tq[NB][4];
    double acc[NB][4];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int ra = tb + 16 * b + ii;
        Aop[b] = kk < 3 ? -2.0 * LT[kk * TT + ra] : 1.0;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = tb + 16 * b + kk + 4 * v;
            const double x = LT[r], y = LT[TT + r], z = LT[2 * TT + r];
            Cop[b][v] = fma(z, z, fma(y, y, x * x));
            tq[b][v] = SS ? (float)A.tq[tile.y + (r < tile.z ? r : 0)] : 0.0f;
            acc[b][v] = 0.0;
        }
    }

    // staged sources: LB[d*TPB + j] = centred coordinate d (d = 3: |s-c|^2),
    // LA[j] = FP64 charge term, LC[j] = (a, b) in FP32.  Entries past n are
    // dummies at the centre whose norm entry is 1e30 and whose charges are
    // zero: every pair with them has r^2 near 1e30 (finite in FP32), adds
    // exactly zero and is never below the threshold, so 16-wide chunks need
    // no tail handling.
    auto put = [&](int n, double x, double y, double z, double a, double bw) {
        const bool real = tid < n;
        const double dx = real ? x - ctr.x : 0.0, dy = real ? y - ctr.y : 0.0, dz = real ? z - ctr.z : 0.0;
        LB[tid] = dx; LB[TPB + tid] = dy; LB[2 * TPB + tid] = dz;
        LB[3 * TPB + tid] = real ? fma(dz, dz, fma(dy, dy, dx * dx)) : 1.0e30;
        a = real ? a : 0.0;
        bw = real ? bw : 0.0;
        LA[tid] = a;
        LC[tid] = make_float2((float)a, (float)bw);
    };
    auto consume = [&](int n, bool check) {
        if (nbv == 0) return;  // no target of this wave in the tile
        if (!check) {
            // two chunks a step: the second chunk's matrix products run while
            // the first chunk's pairs are evaluated (sc <= 224, so both chunks
            // lie within the TPB staged entries; entries past n are dummies)
            for (int sc = 0; sc < n; sc += 32) {
                const int j0 = sc + ii, j1 = sc + 16 + ii;
                const double bop0 = LB[kk * TPB + j0], bop1 = LB[kk * TPB + j1];
                const float2 f0 = LC[j0], f1 = LC[j1];
                const double a0 = LA[j0], a1 = LA[j1];
                v4d D0[NB], D1[NB];
#pragma unroll
                for (int b = 0; b < NB; b++) D0[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(Aop[b], bop0, Cop[b], 0, 0, 0);
#pragma unroll
                for (int b = 0; b < NB; b++) D1[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(Aop[b], bop1, Cop[b], 0, 0, 0);
#pragma unroll
                for (int b = 0; b < NB; b++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        pair_pot32<KER, SS, false>((float)D0[b][v], a0, f0.x, f0.y, tq[b][v], k32, acc[b][v]);
#pragma unroll
                for (int b = 0; b < NB; b++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        pair_pot32<KER, SS, false>((float)D1[b][v], a1, f1.x, f1.y, tq[b][v], k32, acc[b][v]);
            }
            return;
        }
        for (int sc = 0; sc < n; sc += 16) {
            const int j = sc + ii;
            const double bop = LB[kk * TPB + j];
            const float2 f = LC[j];
            const double a64 = LA[j];
            // every row block, valid or not (rows past the tile compute on a
            // clamped target and are never stored): no branches in the chunk
            v4d D[NB];
#pragma unroll
            for (int b = 0; b < NB; b++) D[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(Aop[b], bop, Cop[b], 0, 0, 0);
            if (check) {
                // one vote per chunk; the rare close pairs are fixed before the
                // branch-free evaluation below
                bool close = false;
#pragma unroll
                for (int b = 0; b < NB; b++)
#pragma unroll
                    for (int v = 0; v < 4; v++) close |= D[b][v] < th;
                if (__any(close)) {
#pragma unroll
                    for (int b = 0; b < NB; b++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            if (D[b][v] < th) {
                                const int r = tb + 16 * b + kk + 4 * v;
                                const double ex = LT[r] - LB[j], ey = LT[TT + r] - LB[TPB + j],
                                             ez = LT[2 * TT + r] - LB[2 * TPB + j];
                                const double r2 = fma(ez, ez, fma(ey, ey, ex * ex));
                                // coincident: no contribution (rsq of FLT_MAX and the
                                // exponentials below underflow to zero); the
                                // regularized kernel has no singularity to skip
                                D[b][v] = (KER == KR || r2 > 0.0) ? r2 : 3.4e38;
                            }
                        }
                }
            }
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    pair_pot32<KER, SS, false>((float)D[b][v], a64, f.x, f.y, tq[b][v], k32, acc[b][v]);
        }
    };
    // a list of `total` source points in chunks of TPB; the next chunk's loads
    // are issued before the current chunk is consumed
    auto stream = [&](int total, bool check, auto &&load) {
        double x = 0.0, y = 0.0, z = 0.0, a = 0.0, bw = 0.0;
        if (tid < total) load(tid, x, y, z, a, bw);
        for (int off = 0; off < total; off += TPB) {
            const int n = min(TPB, total - off);
            __syncthreads();
            put(n, x, y, z, a, bw);
            __syncthreads();
            if (off + TPB + tid < total) load(off + TPB + tid, x, y, z, a, bw);
            consume(n, check);
        }
    };

    // source clusters
    const int c0 = A.clu_off[key], c1 = A.clu_off[key + 1];
    for (int c = c0; c < c1; ++c) {
        const int id = A.clu_id[c];
        int P = A.P;
        const double *cx = A.cx, *cy = A.cy, *cz = A.cz, *cq = A.cq, *cw = A.cw;
        size_t base;
        if (A.multi) {
            const int dg = id >> DEG_SHIFT;
            P = (dg + 1) * (dg + 1) * (dg + 1);
            cx = A.dcx[dg]; cy = A.dcy[dg]; cz = A.dcz[dg]; cq = A.dcq[dg]; cw = A.dcw[dg];
            base = (size_t)(id & SLOT_MASK) * P;
        } else {
            base = (size_t)id * P;
        }
        stream(P, can_close && ctr.w < 0.0, [&](int t, double &x, double &y, double &z, double &a, double &bw) {
            const size_t g = base + t;
            x = cx[g]; y = cy[g]; z = cz[g]; a = cq[g];
            if constexpr (SS) bw = cw[g];
        });
    }
    // source particle ranges
    const int s0 = A.seg_off[key], s1 = A.seg_off[key + 1];
    for (int s = s0; s < s1; ++s) {
        const Seg sg = A.segs[s];
        if (A.skip_sym && ((unsigned)sg.y & SEG_SYM_BIT) != 0) continue;  // the symmetric launch's
        // coordinate differences below the threshold (and the skip of
        // coincident pairs, which are closer still: SEG_CLOSE covers them)
        const bool check = can_close && ((unsigned)sg.y & SEG_CLOSE_BIT) != 0;
        const int cnt = (int)((unsigned)sg.y & SEG_COUNT_MASK);
        stream(cnt, check, [&](int t, double &x, double &y, double &z, double &a, double &bw) {
            const int g = sg.x + t;
            x = A.sx[g]; y = A.sy[g]; z = A.sz[g];
            const double q = A.sq[g];
            a = q;
            if constexpr (SS) { bw = A.sw[g]; a = q * bw; }
        });
    }
    // each target's sum: its 16 lanes, butterfly
#pragma unroll
    for (int b = 0; b < NB; b++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double t = acc[b][v];
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) t += __shfl_xor(t, o, 16);
            const int r = tb + 16 * b + kk + 4 * v;
            if (ii == 0 && b < nbv && r < tile.z) {
                const int g = tile.y + r;
                if (A.accumulate) A.out1[g] += t;
                else A.out1[g] = t;
            }
        }
    }
}

template <int KER, bool SS, int NB>
__global__ __launch_bounds__(TPB) void k_near_mfma(InteractArgs A)
{
    constexpr int TT = 4 * 16 * NB;  // targets per pass
    __shared__ double LB[4 * TPB];
    __shared__ double LA[TPB];
    __shared__ float2 LC[TPB];
    __shared__ double LT[3 * TT];
    __shared__ int Lt;
    const int home = blockIdx.x % NQ;
    // one-shot: the block's own tile; persistent: tiles popped from the
    // queues (as k_interact) until every queue is empty
    for (int round = 0;; round++) {
        int t;
        if (!A.persistent) {
            if (round > 0) break;
            t = blockIdx.x;
        } else {
            if (threadIdx.x == 0) {
                int u = -1;
                for (int k = 0; k < NQ && u < 0; k++) {
                    const int q = (home + k) % NQ;
                    const int len = A.qlen[q];
                    if (__hip_atomic_load(&A.cursor[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len)
                        continue;
                    const int j = atomicAdd(&A.cursor[q], 1);
                    if (j < len) u = j * NQ + q;
                }
                Lt = u;
            }
            __syncthreads();
            t = Lt;
            __syncthreads();
            if (t < 0) break;
        }
        const Tile tile = A.tiles[t];
        // a tile larger than one pass runs as consecutive parts
        for (int h = 0; h * TT < tile.z; h++) {
            Tile part = tile;
            part.y = tile.y + h * TT;
            part.z = min(TT, tile.z - h * TT);
            near_mfma_tile<KER, SS, NB>(A, part, LB, LA, LC, LT);
            __syncthreads();
        }
    }
}

__global__ void k_zero_ints(int *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

template <class F> int persistent_blocks(F kernel, int ntiles)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, TPB, 0) != hipSuccess) return -1;
    int nb = std::max(1, cus * std::max(1, per_cu));
    nb = std::min(nb, ntiles);
    return (nb + NQ - 1) / NQ * NQ;
}

template <int KER, bool SS, bool GRID, bool MIXED, int TPT>
hipError_t interact_launch(const InteractArgs &a, hipStream_t st)
{
    auto kern = k_interact<KER, SS, GRID, MIXED, TPT>;
    if (!a.persistent) {
        kern<<<dim3(a.ntiles), dim3(TPB), 0, st>>>(a);
        return hipGetLastError();
    }
    int real = 0;
    for (int q = 0; q < NQ; q++) {
        // the last tile of queue q, slot (qlen-1)*NQ + q, must exist
        if (a.qlen[q] < 0 || (a.qlen[q] > 0 && (long long)(a.qlen[q] - 1) * NQ + q >= a.ntiles))
            return hipErrorInvalidValue;
        real += a.qlen[q];
    }
    if (!a.cursor) return hipErrorInvalidValue;
    const int nb = persistent_blocks(kern, std::max(real, 1));
    if (nb <= 0) return hipErrorInvalidValue;
    k_zero_ints<<<1, 64, 0, st>>>(a.cursor, NQ);
    kern<<<dim3(nb), dim3(TPB), 0, st>>>(a);
    return hipGetLastError();
}

template <int KER, bool SS, int NB>
hipError_t near_mfma_launch(const InteractArgs &a, hipStream_t st)
{
    auto kern = k_near_mfma<KER, SS, NB>;
    if (!a.persistent) {
        kern<<<dim3(a.ntiles), dim3(TPB), 0, st>>>(a);
        return hipGetLastError();
    }
    int real = 0;
    for (int q = 0; q < NQ; q++) {
        if (a.qlen[q] < 0 || (a.qlen[q] > 0 && (long long)(a.qlen[q] - 1) * NQ + q >= a.ntiles))
            return hipErrorInvalidValue;
        real += a.qlen[q];
    }
    if (!a.cursor) return hipErrorInvalidValue;
    const int nb = persistent_blocks(kern, std::max(real, 1));
    if (nb <= 0) return hipErrorInvalidValue;
    k_zero_ints<<<1, 64, 0, st>>>(a.cursor, NQ);
    kern<<<dim3(nb), dim3(TPB), 0, st>>>(a);
    return hipGetLastError();
}

// The kernels beside the Coulomb family: FP64 whatever the precision asked
// (no subtraction on grid targets: the reference has none for them).
template <int KER, bool SS>
hipError_t plain_dispatch(const InteractArgs &a, hipStream_t st)
{
    if (a.grid_targets) {
        if constexpr (SS) return hipErrorInvalidValue;
        else return interact_launch<KER, false, true, false, GRID_TPT>(a, st);
    }
    if (a.tpt == 1) return interact_launch<KER, SS, false, false, 1>(a, st);
    return interact_launch<KER, SS, false, false, 2>(a, st);
}

template <int KER, bool SS>
hipError_t interact_dispatch(const InteractArgs &a, hipStream_t st)
{
    if (a.grid_targets) {
        if (a.precision == 1) return interact_launch<KER, SS, true, true, GRID_TPT>(a, st);
        return interact_launch<KER, SS, true, false, GRID_TPT>(a, st);
    }
    if (a.precision == 1) {
        // mixed particle-target launches form r^2 around the key centres: the
        // skipping kernels on the matrix cores (4 row blocks of 16 targets a
        // wave, 256 targets per pass of a tile), the subtraction kernels on
        // the scaled VALU path
        if (!a.key_ctr) return hipErrorInvalidValue;
        if constexpr (!SS) {
            return near_mfma_launch<KER, SS, 4>(a, st);
        } else {
            if (a.tpt == 1) return interact_launch<KER, SS, false, true, 1>(a, st);
            return interact_launch<KER, SS, false, true, 2>(a, st);
        }
    }
    if (a.tpt == 1) return interact_launch<KER, SS, false, false, 1>(a, st);
    return interact_launch<KER, SS, false, false, 2>(a, st);
}

// ---------------------------------------------------------------------
// Cluster grid coordinates from node tables.
// ---------------------------------------------------------------------
__global__ void k_cluster_points(int numnodes, int nl, const double *tab, double *cx, double *cy,
                                 double *cz)
{
    const int P = nl * nl * nl;
    const size_t total = (size_t)numnodes * P;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (size_t)gridDim.x * blockDim.x) {
        const int node = (int)(g / P), j = (int)(g % P);
        const int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
        const double *tb = tab + (size_t)node * 3 * nl;
        cx[g] = tb[k1]; cy[g] = tb[nl + k2]; cz[g] = tb[2 * nl + k3];
    }
}

// ---------------------------------------------------------------------
// Barycentric weights of the Chebyshev points of the second kind:
// w_j = (-1)^j d_j, d_0 = d_n = 1/2 (clusters.c:446-457).
// ---------------------------------------------------------------------
__device__ __forceinline__ double bary_w(int j, int n)
{
    const double d = (j == 0 || j == n) ? 0.5 : 1.0;
    return (j & 1) ? -d : d;
}

// ---------------------------------------------------------------------
// Upward pass (anterpolation) on the FP64 matrix cores, one block of 256
// threads (4 waves) per output node.  Points (a node's particles, or the
// grid points of its children) arrive in chunks of ACH; for each chunk 3*ACH
// threads compute the per-dimension factors A_d[k][p] = w_k/(x_d - node_d[k])
// (or the exact-hit Kronecker row; clusters.c:506-565, 379-582) into LDS and
// ACH threads the modified charges f1 = q/den (f2 = w/den), den the product
// of the three factor sums.  The grid's charges are then the product
//   Q[(k1,k2), k3] += sum_p (A_x[k1][p] A_y[k2][p]) (A_z[k3][p] f1[p])
// of an NL^2 x ACH matrix and an ACH x NL matrix; with subtraction the
// weights W are the columns NL..2NL-1 of the same product (f2 instead of
// f1).  Tiles of 16x16 go through v_mfma_f64_16x16x4f64 (4 points per
// instruction): wave w owns column tile w % CT and every (4/CT)-th row tile
// (CT = 1 or 2 column tiles), so a step reads each B operand once.
// Operands: lane l supplies A[l%16][l/16] and B[l/16][l%16]; it holds
// D[l/16 + 4v][l%16], v = 0..3.  Out-of-range rows and columns and points
// past the chunk's end supply zeros.
// ---------------------------------------------------------------------
constexpr int ACH = 64;

template <int NL, bool SS>
__global__ __launch_bounds__(256) void k_anterp(AnterpArgs A)
{
    constexpr int P = NL * NL * NL;
    constexpr int R = NL * NL;                       // rows (k1, k2)
    constexpr int C = SS ? 2 * NL : NL;              // columns: k3 (Q), NL + k3 (W)
    constexpr int RT = (R + 15) / 16, CT = (C + 15) / 16;
    static_assert(CT == 1 || CT == 2, "column tiles");
    constexpr int RSTEP = 4 / CT;                    // row tiles between a wave's tiles
    constexpr int TPW = (RT + RSTEP - 1) / RSTEP;    // row tiles per wave (at most)
    constexpr int LDA = ACH + 1;
    constexpr int n = NL - 1;
    __shared__ double LA[3 * NL * LDA];
    __shared__ double Lsum[3 * ACH];
    __shared__ double Lf[2 * ACH];
    __shared__ double Lnode[3 * NL];

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int node = A.nodes[blockIdx.x];
    if (tid < 3 * NL) Lnode[tid] = A.tab[(size_t)node * 3 * NL + tid];

    // this lane's B column: k3 and which charge (0: f1, 1: f2); -1 outside
    const int ct = wv % CT;
    const int bc = ct * 16 + li;
    const int b_k3 = bc < C ? bc % NL : -1, b_f = bc < C ? bc / NL : 0;
    const int rt0 = wv / CT;
    // this lane's A rows, one per row tile of the wave: k1 and k2 (-1 outside)
    int a_k1[TPW], a_k2[TPW];
    v4d acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; t++) {
        const int r = (rt0 + RSTEP * t) * 16 + li;
        const bool ok = rt0 + RSTEP * t < RT && r < R;
        a_k1[t] = ok ? r % NL : -1;
        a_k2[t] = ok ? r / NL : 0;
        acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    }

    // point sets: mode 0 one range of particles (the node's); mode 2 one
    // range of particles (the block's part of a node); mode 1 the grids of
    // all children; mode 3 the grids of the block's children (one entry of
    // the parent-child list), into the entry's row child_off[b]
    const bool kids = A.mode == 1 || A.mode == 3;
    int nsets = 1;
    if (kids) nsets = A.child_off[blockIdx.x + 1] - A.child_off[blockIdx.x];
    const int orow = A.mode == 2 ? (int)blockIdx.x : (A.mode == 3 ? A.child_off[blockIdx.x] : node);
    for (int set = 0; set < nsets; ++set) {
        int pbeg = 0, pcnt = P, child = -1;
        if (A.mode == 0) {
            pbeg = A.ibeg[node];
            pcnt = A.iend[node] - pbeg;
        } else if (A.mode == 2) {
            pbeg = A.ibeg[blockIdx.x];
            pcnt = A.iend[blockIdx.x] - pbeg;
        } else {
            child = A.child_id[A.child_off[blockIdx.x] + set];
        }
        for (int c0 = 0; c0 < pcnt; c0 += ACH) {
            __syncthreads();  // the previous chunk's factors have been read
            if (tid < 3 * ACH) {
                const int p = tid % ACH, d = tid / ACH;
                const bool valid = c0 + p < pcnt;
                double coord = 0.0;
                if (valid) {
                    if (!kids) {
                        const int g = pbeg + c0 + p;
                        coord = d == 0 ? A.px[g] : (d == 1 ? A.py[g] : A.pz[g]);
                    } else {
                        const int jp = c0 + p;
                        const int kd = d == 0 ? jp % NL : (d == 1 ? (jp / NL) % NL : jp / (NL * NL));
                        coord = A.tab[(size_t)child * 3 * NL + d * NL + kd];
                    }
                }
                double sum = 0.0;
                int hit = -1;
#pragma unroll
                for (int k = 0; k < NL; k++) {
                    const double c = coord - Lnode[d * NL + k];
                    if (fabs(c) < DBL_MIN) hit = k;
                    const double v = bary_w(k, n) / c;
                    sum += v;
                    LA[(d * NL + k) * LDA + p] = v;
                }
                if (hit >= 0 || !valid) {
#pragma unroll
                    for (int k = 0; k < NL; k++) LA[(d * NL + k) * LDA + p] = (valid && k == hit) ? 1.0 : 0.0;
                    sum = 1.0;
                }
                Lsum[d * ACH + p] = sum;
            }
            __syncthreads();
            if (tid < ACH) {
                const int p = tid;
                double f1 = 0.0, f2 = 0.0;
                if (c0 + p < pcnt) {
                    double den = 1.0;
                    den *= Lsum[p];
                    den *= Lsum[ACH + p];
                    den *= Lsum[2 * ACH + p];
                    if (!kids) {
                        const int g = pbeg + c0 + p;
                        const double q = A.pq[g];
                        if (SS) { const double w = A.pw[g]; f1 = q * w; f2 = w; }
                        else f1 = q;
                    } else {
                        const size_t g = (size_t)child * P + c0 + p;
                        f1 = A.Q[g];
                        if (SS) f2 = A.W[g];
                    }
                    f1 /= den;
                    if (SS) f2 /= den;
                }
                Lf[p] = f1;
                Lf[ACH + p] = f2;
            }
            __syncthreads();
            // points past the end have zero factors and charges
            const int np4 = min(ACH, (pcnt - c0 + 3) & ~3);
            for (int p0 = 0; p0 < np4; p0 += 4) {
                const int p = p0 + lk;
                const double bop = b_k3 >= 0 ? LA[(2 * NL + b_k3) * LDA + p] * Lf[b_f * ACH + p] : 0.0;
#pragma unroll
                for (int t = 0; t < TPW; t++) {
                    if (rt0 + RSTEP * t >= RT) continue;  // wave-uniform
                    const double aop = a_k1[t] >= 0 ? LA[a_k1[t] * LDA + p] * LA[(NL + a_k2[t]) * LDA + p] : 0.0;
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, bop, acc[t], 0, 0, 0);
                }
            }
        }
    }
    // lane l holds rows (k1, k2) = r, r = rt*16 + l/16 + 4v, of column bc
    if (b_k3 < 0) return;
    double *out = A.mode == 3 ? (b_f == 0 ? A.Qo : A.Wo) : (b_f == 0 ? A.Q : A.W);
#pragma unroll
    for (int t = 0; t < TPW; t++) {
        if (rt0 + RSTEP * t >= RT) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = (rt0 + RSTEP * t) * 16 + lk + 4 * v;
            if (r < R) out[(size_t)orow * P + r + R * b_k3] = acc[t][v];
        }
    }
}

// ---------------------------------------------------------------------
// Downward pass: evaluate the interpolant of a node's grid values at points
// (the child's grid points, or the leaf's targets).  Per point the factors
// w_k/(x_d - node_d[k]) live in registers (NL is a template parameter);
// the value is  D * sum_k3 wz sum_k2 wy sum_k1 wx Q[k1,k2,k3]  with
// D = ((1/sx)/sy)/sz over the dimensions without an exact hit
// (interaction_compute_downpass.c:210-303, 395-479).
// ---------------------------------------------------------------------
template <int NL>
__device__ __forceinline__ void point_factors(double x, const double *nodes, double *wv, double &den,
                                              bool &hit_any)
{
    double sum = 0.0;
    int hit = -1;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const double c = x - nodes[k];
        if (fabs(c) < DBL_MIN) hit = k;
        const double v = bary_w(k, NL - 1) / c;
        wv[k] = v;
        sum += v;
    }
    if (hit >= 0) {
#pragma unroll
        for (int k = 0; k < NL; k++) wv[k] = k == hit ? 1.0 : 0.0;
    } else {
        den /= sum;
    }
    hit_any = hit_any || hit >= 0;
}

constexpr int DTPB = 256;  // threads per block of the downward pass (4 waves)
constexpr int DCH = 64;    // points per chunk (16 per wave)

// Downward pass on the FP64 matrix cores.  For a chunk of points (a
// child's grid points, or a tile of a leaf's targets) 3*DCH threads put the
// per-dimension factors into LDS (point_factors: exact hits become Kronecker
// rows, den collects 1/sum over the other dimensions).  Wave w takes points
// 16w..16w+15 of the chunk and forms
//   T[p][(k2,k3)] = sum_k1 wx[p][k1] Q[k1 + NL (k2 + NL k3)]
// in 16x16 tiles with v_mfma_f64_16x16x4f64 (K = k1, padded to a multiple
// of 4 with zeros), then each lane weights its entries of T by
// wy[p][k2] wz[p][k3] and the 16 lanes of a row add them (DPP row_shr
// scan); the value at p is den * that sum
// (interaction_compute_downpass.c:210-303, 395-479; only the order of the
// sums differs).
template <int NL, bool SS>
__global__ __launch_bounds__(DTPB) void k_down(DownArgs A)
{
    constexpr int P = NL * NL * NL;
    constexpr int N2 = NL * NL;                      // columns (k2, k3)
    constexpr int CT = (N2 + 15) / 16;               // column tiles
    constexpr int KS = (NL + 3) / 4;                 // k1 steps of 4
    constexpr int LDW = DCH + 1;
    __shared__ double LQ[P];
    __shared__ double LW[SS ? P : 1];
    __shared__ double Lnode[3 * NL];
    __shared__ double Lw[3 * NL * LDW];              // factors [d][k][p]
    __shared__ double Lsum[3 * DCH];                 // 1/sum of a dimension (1 on a hit)
    __shared__ double Lv[2 * DCH];                   // the chunk's results (Q, W)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    int src_node, pbeg = 0, pcnt = P, dst_node = -1;
    if (A.mode == 0) {
        dst_node = A.nodes[blockIdx.x];
        src_node = A.parent[dst_node];
    } else {
        const Tile t = A.tiles[blockIdx.x];
        src_node = t.x; pbeg = t.y; pcnt = t.z;
    }
    for (int j = tid; j < P; j += DTPB) {
        LQ[j] = A.Q[(size_t)src_node * P + j];
        if (SS) LW[j] = A.W[(size_t)src_node * P + j];
    }
    if (tid < 3 * NL) Lnode[tid] = A.tab[(size_t)src_node * 3 * NL + tid];
    for (int c0 = 0; c0 < pcnt; c0 += DCH) {
        __syncthreads();  // LQ, Lnode ready; the previous chunk's factors read
        if (tid < 3 * DCH) {
            const int p = tid % DCH, d = tid / DCH, i = c0 + p;
            const bool valid = i < pcnt;
            double x = 0.0;
            if (valid) {
                if (A.mode == 0) {
                    const int kd = d == 0 ? i % NL : (d == 1 ? (i / NL) % NL : i / (NL * NL));
                    x = A.tab[(size_t)dst_node * 3 * NL + d * NL + kd];
                } else {
                    const int g = pbeg + i;
                    x = d == 0 ? A.tx[g] : (d == 1 ? A.ty[g] : A.tz[g]);
                }
            }
            double wk[NL];
            double den = 1.0;
            bool h = false;
            point_factors<NL>(x, Lnode + d * NL, wk, den, h);
#pragma unroll
            for (int k = 0; k < NL; k++) Lw[(d * NL + k) * LDW + p] = valid ? wk[k] : 0.0;
            Lsum[d * DCH + p] = den;
        }
        __syncthreads();
        // stage 1 and 2 for points pw + 4v + lk (rows of the tile), pw = 16 wv
        const int pw = 16 * wv;
        double sq[4] = {0.0, 0.0, 0.0, 0.0}, sw[4] = {0.0, 0.0, 0.0, 0.0};
        if (c0 + pw < pcnt) {  // wave-uniform: the wave has points
#pragma unroll 1
            for (int ct = 0; ct < CT; ct++) {
                const int col = ct * 16 + li;             // B column / this lane's column of T
                const bool cok = col < N2;
                v4d dq = {0.0, 0.0, 0.0, 0.0}, dw = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < KS; ks++) {
                    const int kA = ks * 4 + lk;           // k1 of this lane's A and B entries
                    const double aop = kA < NL ? Lw[kA * LDW + pw + li] : 0.0;
                    const double bq = (kA < NL && cok) ? LQ[kA + NL * col] : 0.0;
                    dq = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, bq, dq, 0, 0, 0);
                    if (SS) {
                        const double bw = (kA < NL && cok) ? LW[kA + NL * col] : 0.0;
                        dw = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, bw, dw, 0, 0, 0);
                    }
                }
                if (cok) {
                    const int k2 = col % NL, k3 = col / NL;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int p = pw + lk + 4 * v;
                        const double f = Lw[(NL + k2) * LDW + p] * Lw[(2 * NL + k3) * LDW + p];
                        sq[v] = fma(f, dq[v], sq[v]);
                        if (SS) sw[v] = fma(f, dw[v], sw[v]);
                    }
                }
            }
            // the 16 lanes of a row: lane 15 of the row ends with the sum
#pragma unroll
            for (int v = 0; v < 4; v++) {
                sq[v] = group_sum64<16>(sq[v]);
                if (SS) sw[v] = group_sum64<16>(sw[v]);
                if (li == 15) {
                    const int p = pw + lk + 4 * v;
                    const double den = Lsum[p] * Lsum[DCH + p] * Lsum[2 * DCH + p];
                    Lv[p] = den * sq[v];
                    if (SS) Lv[DCH + p] = den * sw[v];
                }
            }
        }
        __syncthreads();
        if (tid < DCH && c0 + tid < pcnt) {
            const int i = c0 + tid;
            const double vq = Lv[tid], vw = SS ? Lv[DCH + tid] : 0.0;
            if (A.mode == 0) {
                const size_t o = (size_t)dst_node * P + i;
                A.Q[o] += vq;
                if (SS) A.W[o] += vw;
            } else {
                const int g = pbeg + i;
                // subtraction: interpolate Q - tq*W (downpass.c:836)
                A.pot[g] += SS ? fma(-A.tq[g], vw, vq) : vq;
            }
        }
    }
}

template <bool SS>
hipError_t anterp_dispatch(const AnterpArgs &a, hipStream_t st)
{
    dim3 g(a.nblocks), b(256);
    switch (a.degree) {
#define CASE(D) case D: k_anterp<D + 1, SS><<<g, b, 0, st>>>(a); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
#undef CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Q[d*P + j] = sum over parts k in [off[s], off[s+1]) of Qp[k*P + j] (and W),
// in part order; d = dst[s], or s without dst
__global__ void k_sum_parts(int nslots, int P, const int *off, const double *Qp, const double *Wp,
                            double *Q, double *W, const int *dst)
{
    const size_t total = (size_t)nslots * P;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(g / P), j = (int)(g % P);
        double q = 0.0, w = 0.0;
        for (int k = off[s]; k < off[s + 1]; k++) {
            q += Qp[(size_t)k * P + j];
            if (W) w += Wp[(size_t)k * P + j];
        }
        const size_t o = dst ? (size_t)dst[s] * P + j : g;
        Q[o] = q;
        if (W) W[o] = w;
    }
}

__global__ void k_sum_rows(int nslots, int P, const int *off, const long long *row, const double *R,
                           const double *RW, double *Q, double *W)
{
    const size_t total = (size_t)nslots * P;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(g / P), j = (int)(g % P);
        double q = 0.0, w = 0.0;
        for (int k = off[s]; k < off[s + 1]; k++) {
            q += R[row[k] + j];
            if (W) w += RW[row[k] + j];
        }
        Q[g] = q;
        if (W) W[g] = w;
    }
}

template <bool SS>
hipError_t down_dispatch(const DownArgs &a, hipStream_t st)
{
    dim3 g(a.nblocks), b(DTPB);
    switch (a.degree) {
#define CASE(D) case D: k_down<D + 1, SS><<<g, b, 0, st>>>(a); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
#undef CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void k_finish(int n, const double *pot, const double *tq, double ss, const int *order, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[order[i]] = ss != 0.0 ? fma(ss, tq[i], pot[i]) : pot[i];
}

// Direct sum: one target per thread; blockIdx.y selects a chunk of the
// sources, so the sum over all N sources spreads over many blocks.  Partial
// sums go to part[chunk][m] and k_direct_reduce adds them in chunk order
// (deterministic), then the subtraction correction.
template <int KER, bool SS>
__global__ __launch_bounds__(256) void k_direct(KParams kp, int m, const double *tx, const double *ty,
                                                const double *tz, const double *tq, int n, int chunk,
                                                const double *sx, const double *sy, const double *sz,
                                                const double *sq, const double *sw, double *part)
{
    __shared__ double4 Ls[256];
    __shared__ double Lb[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ic = i < m ? i : 0;
    const double x = tx[ic], y = ty[ic], z = tz[ic], q = SS ? tq[ic] : 0.0;
    const int s0 = blockIdx.y * chunk, s1 = min(n, s0 + chunk);
    double acc = 0.0;
    for (int off = s0; off < s1; off += 256) {
        const int cnt = min(256, s1 - off);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const int g = off + threadIdx.x;
            double a = sq[g], b = 0.0;
            if (SS) { b = sw[g]; a = a * b; }
            Ls[threadIdx.x] = make_double4(sx[g], sy[g], sz[g], a);
            Lb[threadIdx.x] = b;
        }
        __syncthreads();
        for (int j = 0; j < cnt; ++j) {
            const double4 s = Ls[j];
            pair_pot<KER, SS, true>(x - s.x, y - s.y, z - s.z, s.w, Lb[j], q, kp, acc);
        }
    }
    if constexpr (KER == KAT) acc *= kp.c2;
    if (i < m) part[(size_t)blockIdx.y * m + i] = acc;
}

__global__ void k_direct_reduce(int m, int nchunks, const double *part, const double *tq, double ss,
                                double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double acc = 0.0;
    for (int c = 0; c < nchunks; c++) acc += part[(size_t)c * m + i];
    out[i] = ss != 0.0 ? fma(ss, tq[i], acc) : acc;
}

}  // namespace

hipError_t launch_interact(const InteractArgs &a, hipStream_t st)
{
    if (a.ntiles <= 0) return hipSuccess;
    const bool ss = a.singularity == 2;
    switch (a.kernel) {
        case KC: return ss ? interact_dispatch<KC, true>(a, st) : interact_dispatch<KC, false>(a, st);
        case KY: return ss ? interact_dispatch<KY, true>(a, st) : interact_dispatch<KY, false>(a, st);
        case KR: return ss ? interact_dispatch<KR, true>(a, st) : interact_dispatch<KR, false>(a, st);
        case KRY: return ss ? plain_dispatch<KRY, true>(a, st) : plain_dispatch<KRY, false>(a, st);
        case KAT: return ss ? hipErrorInvalidValue : plain_dispatch<KAT, false>(a, st);
        case KSR: return ss ? hipErrorInvalidValue : plain_dispatch<KSR, false>(a, st);
        case KMQ: return ss ? hipErrorInvalidValue : plain_dispatch<KMQ, false>(a, st);
        case KRU: return ss ? hipErrorInvalidValue : plain_dispatch<KRU, false>(a, st);
        case KRV: return ss ? hipErrorInvalidValue : plain_dispatch<KRV, false>(a, st);
        case KUS: return ss ? hipErrorInvalidValue : plain_dispatch<KUS, false>(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_grid_pencil(const InteractArgs &a, hipStream_t st)
{
    if (a.ntiles <= 0) return hipSuccess;
    const bool ss = a.singularity == 2;
    if (a.precision == 0) {
        switch (a.kernel) {
#define L(K, S) k_grid_pencil64<K, S><<<dim3(a.ntiles), dim3(256), 0, st>>>(a)
            case KC: if (ss) L(KC, true); else L(KC, false); break;
            case KY: if (ss) L(KY, true); else L(KY, false); break;
            case KR: if (ss) return hipErrorInvalidValue; L(KR, false); break;
#undef L
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (a.kernel) {
#define L(K, S) k_grid_pencil<K, S><<<dim3(a.ntiles), dim3(256), 0, st>>>(a)
        case KC: if (ss) L(KC, true); else L(KC, false); break;
        case KY: if (ss) L(KY, true); else L(KY, false); break;
        case KR: if (ss) return hipErrorInvalidValue; L(KR, false); break;
#undef L
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_near_sym(const SymArgs &a, hipStream_t st)
{
    if (a.nitems <= 0) return hipSuccess;
    const bool ss = a.singularity == 2;
    if (a.fp64) {
        switch (a.kernel) {
#define L(K, S) k_near_sym64<K, S, SYM64_TPT, SYM64_G, SYM64_NW><<<dim3(a.nitems), dim3(64 * SYM64_NW), 0, st>>>(a)
            case KC: if (ss) L(KC, true); else L(KC, false); break;
            case KY: if (ss) L(KY, true); else L(KY, false); break;
            case KR: if (ss) return hipErrorInvalidValue; L(KR, false); break;
#undef L
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (a.kernel) {
#define L(K, S) k_near_sym<K, S, sym_tpt(K, S), SYM_G, SYM_NW, sym_wpe(K, S)><<<dim3(a.nitems), dim3(64 * SYM_NW), 0, st>>>(a)
        case KC: if (ss) L(KC, true); else L(KC, false); break;
        case KY: if (ss) L(KY, true); else L(KY, false); break;
        case KR: if (ss) return hipErrorInvalidValue; L(KR, false); break;
#undef L
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sym_reduce(const SymReduceArgs &a, hipStream_t st)
{
    if (a.ntiles <= 0) return hipSuccess;
    k_sym_reduce<<<dim3(a.ntiles), dim3(TPB_SYM_REDUCE), 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_sum_parts(int nslots, int P, const int *off, const double *Qp, const double *Wp,
                            double *Q, double *W, hipStream_t st, const int *dst)
{
    const size_t total = (size_t)nslots * P;
    if (total == 0) return hipSuccess;
    int blocks = (int)std::min<size_t>((total + 255) / 256, 65535);
    k_sum_parts<<<blocks, 256, 0, st>>>(nslots, P, off, Qp, Wp, Q, W, dst);
    return hipGetLastError();
}

hipError_t launch_sum_rows(int nslots, int P, const int *off, const long long *row, const double *R,
                           const double *RW, double *Q, double *W, hipStream_t st)
{
    const size_t total = (size_t)nslots * P;
    if (total == 0) return hipSuccess;
    int blocks = (int)std::min<size_t>((total + 255) / 256, 65535);
    k_sum_rows<<<blocks, 256, 0, st>>>(nslots, P, off, row, R, RW, Q, W);
    return hipGetLastError();
}

hipError_t launch_cluster_points(int numnodes, int nl, const double *tab, double *cx, double *cy,
                                 double *cz, hipStream_t st)
{
    const size_t total = (size_t)numnodes * nl * nl * nl;
    if (total == 0) return hipSuccess;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 65535) blocks = 65535;
    k_cluster_points<<<blocks, 256, 0, st>>>(numnodes, nl, tab, cx, cy, cz);
    return hipGetLastError();
}

hipError_t launch_anterp(const AnterpArgs &a, hipStream_t st)
{
    if (a.nblocks <= 0) return hipSuccess;
    return a.subtraction ? anterp_dispatch<true>(a, st) : anterp_dispatch<false>(a, st);
}

hipError_t launch_down(const DownArgs &a, hipStream_t st)
{
    if (a.nblocks <= 0) return hipSuccess;
    return a.subtraction ? down_dispatch<true>(a, st) : down_dispatch<false>(a, st);
}

hipError_t launch_finish(int n, const double *pot, const double *tq, double ss_param, const int *order,
                         double *out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_finish<<<(n + 255) / 256, 256, 0, st>>>(n, pot, tq, ss_param, order, out);
    return hipGetLastError();
}

hipError_t launch_direct(int kernel, int singularity, KParams kp, int m, const double *tx,
                         const double *ty, const double *tz, const double *tq, int n,
                         const double *sx, const double *sy, const double *sz, const double *sq,
                         const double *sw, double *part, int nchunks, double *out, hipStream_t st)
{
    if (m <= 0 || n <= 0 || nchunks <= 0) return hipSuccess;
    const int chunk = (n + nchunks - 1) / nchunks;
    dim3 g((m + 255) / 256, nchunks), b(256);
    const bool ss = singularity == 2;
#define L(K, S) k_direct<K, S><<<g, b, 0, st>>>(kp, m, tx, ty, tz, tq, n, chunk, sx, sy, sz, sq, sw, part)
    switch (kernel) {
        case KC: if (ss) L(KC, true); else L(KC, false); break;
        case KY: if (ss) L(KY, true); else L(KY, false); break;
        case KR: if (ss) L(KR, true); else L(KR, false); break;
        case KRY: if (ss) L(KRY, true); else L(KRY, false); break;
        case KAT: L(KAT, false); break;
        case KSR: L(KSR, false); break;
        case KMQ: L(KMQ, false); break;
        case KRU: L(KRU, false); break;
        case KRV: L(KRV, false); break;
        case KUS: L(KUS, false); break;
        default: return hipErrorInvalidValue;
    }
#undef L
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_direct_reduce<<<(m + 255) / 256, 256, 0, st>>>(m, nchunks, part, tq, ss ? kp.ss_param : 0.0, out);
    return hipGetLastError();
}

}  // namespace btk
