// Device-side building blocks of the Fock build: QUICK-compatible Boys
// function, per-primitive-quartet VRR parameters, and the two kernel
// templates (Fock digestion and Schwarz diagonal) that every generated
// (La Lb | Lc Ld) class instantiates.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "gen/boys_grid.h"
#include "qf_types.h"

namespace qf {

// Pointers fetched from descriptors are generic; these casts tell the
// compiler which memory they address, so wave-uniform reads become scalar
// loads (constant address space) and per-lane reads / atomics use the
// global-memory instructions instead of flat ones.  (The host pass of the
// HIP compiler sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define QF_AS_CONST __attribute__((address_space(4)))
#define QF_AS_GLOBAL __attribute__((address_space(1)))
#define QF_AS_LDS __attribute__((address_space(3)))
#else
#define QF_AS_CONST
#define QF_AS_GLOBAL
#define QF_AS_LDS
#endif
template <typename T>
__host__ __device__ __forceinline__ const QF_AS_CONST T* cptr(const T* p)
{
    return (const QF_AS_CONST T*)(p);
}
template <typename T>
__host__ __device__ __forceinline__ const QF_AS_GLOBAL T* gptr(const T* p)
{
    return (const QF_AS_GLOBAL T*)(p);
}
template <typename T>
__device__ __forceinline__ QF_AS_GLOBAL T* gptr_rw(T* p)
{
    return (QF_AS_GLOBAL T*)(p);
}

// A value that the caller knows to be the same in every active lane, moved
// to a scalar register.  Loop counters of loops whose exit differs between
// lanes are kept in vector registers by the compiler (it cannot tell that
// the active lanes agree), and so would be the loads they index.
__host__ __device__ __forceinline__ int wave_uniform(int x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(x);
#else
    return x;
#endif
}

// x of lane `src` (a wave-uniform index) in every lane, as a scalar value
__device__ __forceinline__ double wave_read_lane(double x, int src)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(x): the hardware reciprocal square root on the device, the C
// library on the host (the host path only serves CPU checks of this code).
__host__ __device__ __forceinline__ double qf_rsqrt(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrt(x);
#else
    return 1.0 / std::sqrt(x);
#endif
}

// 1/x: the hardware reciprocal refined by two Newton steps on the device
// (within an ulp of the division; x = 0 gives NaN, which no caller uses),
// the division on the host.
__host__ __device__ __forceinline__ double qf_rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}

// 1/sqrt(x) for normal x > 0: the hardware reciprocal square root refined
// by two Newton steps r += r (1 - x r^2) / 2 (within an ulp); x = 0 gives
// NaN, which only the branches that do not use it see.
__host__ __device__ __forceinline__ double qf_rsqrt_fine(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rsq(x);
    r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
    return fma(0.5 * r, fma(-x * r, r, 1.0), r);
#else
    return 1.0 / std::sqrt(x);
#endif
}

// e^-x for x >= 0.  Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, with
// ln2 split into a high part whose product with n is exact (|n| < 2^20) and
// a low part (the fdlibm split); e^-r from its Taylor series to r^13
// (truncation below 5e-18 relative at |r| = ln2/2); then 2^-n by ldexp.
// x is clamped to 750, where e^-x already rounds to zero.
__host__ __device__ __forceinline__ double qf_exp_neg(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    x = fmin(x, 750.0);
    const double n = rint(x * 1.44269504088896340736);
    const double r = fma(-n, 1.90821492927058770002e-10, fma(-n, 6.93147180369123816490e-01, x));
    const double t = -r;
    double p = 1.6059043836821614599e-10;      // 1/13!
    p = fma(p, t, 2.0876756987868098979e-09);  // 1/12!
    p = fma(p, t, 2.5052108385441718775e-08);  // 1/11!
    p = fma(p, t, 2.7557319223985890653e-07);  // 1/10!
    p = fma(p, t, 2.7557319223985890653e-06);  // 1/9!
    p = fma(p, t, 2.4801587301587301587e-05);  // 1/8!
    p = fma(p, t, 1.9841269841269841270e-04);  // 1/7!
    p = fma(p, t, 1.3888888888888888889e-03);  // 1/6!
    p = fma(p, t, 8.3333333333333333333e-03);  // 1/5!
    p = fma(p, t, 4.1666666666666666667e-02);  // 1/4!
    p = fma(p, t, 1.6666666666666666667e-01);  // 1/3!
    p = fma(p, t, 0.5);
    p = fma(p, t, 1.0);
    p = fma(p, t, 1.0);
    return ldexp(p, -(int)n);
#else
    return std::exp(-x);
#endif
}

template <typename R>
struct VrrPar {
    R PA[3], WP[3], QC[3], WQ[3];
    R oo2z, oo2n, oo2zn, rz, rn;   // 1/2z, 1/2n, 1/2(z+n), rho/z, rho/n
};

// ---------------------------------------------------------------------------
// Boys function F_m(X), m = 0..L, evaluated exactly as QUICK's FmT
// (src/subs/fmt.f90:13-85, src/gpu/gpu_fmt.h): the same piecewise fits for
// F_0, the same branch limits, upward recursion above X = 0.1 (or 1e-4 when
// maxm < 4), otherwise downward recursion from (1-X)/(2 maxm+1).  maxm is
// QUICK's NABCD: the sum of the four *original* shells' highest L (an SP
// shell counts 1 even for its s part), which selects the small-X branch.
// The fit coefficients are QUICK's; parity with QUICK needs them.
// ---------------------------------------------------------------------------
// QUICK's fits as one table: interval k (X <= 1, 1..3, 3..5, 5..10, 10..15,
// 15..33, > 33), Horner coefficients of the fitted polynomial, highest power
// first, padded in front with zeros (fma(0,u,0) = 0 and fma(0,u,c) = c, so
// the padded evaluation is bit-identical to the unpadded one).  Intervals
// 0-2: F1 = p(u), u = X, X-2, X-4, and F_0 = 2X F1 + e^-X; intervals 3-6:
// u = 1/X and F_0 = p(u) e^-X + sqrt(pi/4/X).
// After the rows: the interval of X by ceil(X) - 1 (clamped to 0..33), as
// a double k, so that the interval is one table read instead of six
// comparisons: X <= 1 -> 0, (1, 3] -> 1, (3, 5] -> 2, (5, 10] -> 3,
// (10, 15] -> 4, (15, 33] -> 5, > 33 -> 6 (X > b  <=>  ceil(X) - 1 >= b for
// the integer bounds b).
constexpr int kBoysRows = 7, kBoysCols = 12, kBoysMap = kBoysRows * kBoysCols, kBoysMapN = 34;
constexpr int kBoysTab = kBoysMap + kBoysMapN;
__host__ __device__ constexpr double kBoysFit[kBoysTab] = {
    0.0, 0.0, -8.36313918003957e-08, 1.21222603512827e-06, -1.15662609053481e-05, 9.25197374512647e-05,
    -6.40994113129432e-04, 3.78787044215009e-03, -1.85185172458485e-02, 7.14285713298222e-02,
    -1.99999999997023e-01, 3.33333333333318e-01,
    -1.61702782425558e-10, 1.96215250865776e-09, -2.14234468198419e-08, 2.17216556336318e-07,
    -1.98850171329371e-06, 1.62429321438911e-05, -1.16740298039895e-04, 7.24888732052332e-04,
    -3.79490003707156e-03, 1.61723488664661e-02, -5.29428148329736e-02, 1.15702180856167e-01,
    -2.62453564772299e-11, 3.24031041623823e-10, -3.614965656163e-09, 3.760256799971e-08,
    -3.553558319675e-07, 3.022556449731e-06, -2.290098979647e-05, 1.526537461148e-04,
    -8.81947375894379e-04, 4.33207949514611e-03, -1.75257821619926e-02, 5.28406320615584e-02,
    0.0, 0.0, 0.0, 0.0, 0.0, 4.6897511375022e-01, -6.9955602298985e-01, 5.3689283271887e-01,
    -3.2883030418398e-01, 2.4645596956002e-01, -4.9984072848436e-01, -3.1501078774085e-06,
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.8784686463512e-01, 2.2991849164985e-01,
    -4.9893752514047e-01, -2.1916512131607e-05,
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.9623264149430e-01, -4.9695241464490e-01,
    -6.0156581186481e-05,
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
    // interval map, ceil(X) - 1 = 0 .. 33
    0, 1, 1, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 6};

// The kernels read the table from LDS: per-lane reads of a tiny table, and
// no fit constant is held in a register across the primitive loops.
#define QF_BOYS_LDS_DECL __shared__ double qf_boys_tab[kBoysTab]
#define QF_BOYS_LDS_FILL()                                              \
    do {                                                                \
        for (int _x = threadIdx.x; _x < kBoysTab; _x += blockDim.x)     \
            qf_boys_tab[_x] = kBoysFit[_x];                             \
        __syncthreads();                                                \
    } while (0)
#define QF_BOYS_LDS_PTR ((const QF_AS_LDS double*)qf_boys_tab)

// EXTRA: how far maxm may exceed L on the downward branch (L + EXTRA + 1 at
// most): 3 for the Fock classes (the s parts of SP shells), 4 for the
// gradient, whose maxm carries QUICK's extra +1 (eri_grad.fh:104-106).
template <int L, typename Tab, int EXTRA = 3>
__host__ __device__ __forceinline__ void boys_quick(const double X, const int maxm, double (&F)[L + 1], Tab tab)
{
    // 1/sqrt(X) once: 1/X is its square and sqrt(pi/4X) a multiple of it
    const double RS = qf_rsqrt_fine(X);
    const double XINV = RS * RS;
    const bool up = (X > 1.0e-1) || (X > 1.0e-4 && maxm < 4);
    // interval index from the map (no branches); both forms of F0 evaluated
    // and one selected (measured on (H2O)32: 7.6 % faster than branching and
    // the library exp / sqrt)
    const double E = qf_exp_neg(X);
    const int ci = (int)fmin(fmax(ceil(X) - 1.0, 0.0), double(kBoysMapN - 1));
    const double kd = tab[kBoysMap + ci];
    const int k = (int)kd;
    const double u = kd >= 3.0 ? XINV : fma(-2.0, kd, X);
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < kBoysCols; ++c) p = fma(p, u, tab[k * kBoysCols + c]);
    const double WW1 = k >= 3 ? fma(p, E, 0.88622692545275801365 * RS) : fma(X + X, p, E);
    if (up) {
        F[0] = WW1;
        const double h = 0.5 * XINV;
#pragma unroll
        for (int m = 1; m <= L; ++m) F[m] = ((2 * m - 1) * F[m - 1] - E) * h;
    } else {
        // maxm exceeds L only by the s parts of SP shells (at most 4)
        const double twot = X * 2.0;
        double f = (1.0 - X) * qf_rcp(2 * maxm + 1.0);
#pragma unroll
        for (int m = L + EXTRA; m >= L; --m)
            if (m < maxm) f = (twot * f + E) * (1.0 / (2 * m + 1));
        F[L] = f;
#pragma unroll
        for (int m = L - 1; m >= 0; --m) F[m] = (twot * F[m + 1] + E) * (1.0 / (2 * m + 1));
    }
}

// e^-x and 1/sqrt(x) in FP32: the hardware exp2 and reciprocal square root
// on the device, the C library on the host (CPU checks only).
__host__ __device__ __forceinline__ float qf_expf_neg(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(-x);
#else
    return std::exp(-x);
#endif
}

__host__ __device__ __forceinline__ float qf_rsqrtf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / std::sqrt(x);
#endif
}

// ---------------------------------------------------------------------------
// Boys function F_m(X), m = 0..L, in FP32, for the FP32 quartets of
// split-precision builds (their integrals are below the split threshold,
// so FP32 accuracy, ~1e-7 relative, is what they need; QUICK's FmT
// branches and fits serve the FP64 quartets).  Inside the grid (X < 30):
// F_L by a 5-term Taylor series about the nearest grid point of kBoysGrid
// (dF_m/dX = -F_{m+1}), truncation below 5e-8 relative; above it the
// asymptotic F_L = sqrt(pi/4X) prod (2m-1)/2X (e^-X < 1e-13).  Then the
// downward recursion F_m = (2X F_{m+1} + e^-X) / (2m+1), which is stable.
// ---------------------------------------------------------------------------
template <int L, typename GT>
__host__ __device__ __forceinline__ void boys_f32(const float X, float (&F)[L + 1], GT grid)
{
    static_assert(L + 4 < kBoysGridM, "Boys grid table too narrow for this class");
    constexpr float kXmax = (kBoysGridN - 1) * kBoysGridH;
    const float E = qf_expf_neg(X);
    const int i = X < kXmax ? int(fmaf(X, kBoysGridInvH, 0.5f)) : kBoysGridN - 1;
    const float d = float(i) * kBoysGridH - X;
    const auto row = grid + i * kBoysGridM + L;
    float f = row[4];
    f = fmaf(f, d * 0.25f, row[3]);
    f = fmaf(f, d * (1.0f / 3.0f), row[2]);
    f = fmaf(f, d * 0.5f, row[1]);
    f = fmaf(f, d, row[0]);
    const float rx = qf_rsqrtf(X);
    float a = 0.88622692545275801365f * rx;   // sqrt(pi/4X)
    const float h = 0.5f * rx * rx;
#pragma unroll
    for (int m = 1; m <= L; ++m) a *= float(2 * m - 1) * h;
    F[L] = X < kXmax ? f : a;
    const float twox = X + X;
#pragma unroll
    for (int m = L - 1; m >= 0; --m) F[m] = fmaf(twox, F[m + 1], E) * (1.0f / float(2 * m + 1));
}

// cons factors of QUICK's Cartesian components (basis.f90:567-720): d
// xy, xz, yz sqrt(3); f xxy, xyy, xxz, yyz, xzz, yzz sqrt(5), xyz sqrt(15)
__host__ __device__ constexpr double kCons[4][10] = {
    {1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {1.0, 1.0, 1.0, 0, 0, 0, 0, 0, 0, 0},
    {1.0, 1.7320508075688772935, 1.0, 1.7320508075688772935, 1.7320508075688772935, 1.0, 0, 0, 0, 0},
    {1.0, 2.2360679774997896964, 2.2360679774997896964, 1.0, 2.2360679774997896964, 3.8729833462074168852,
     2.2360679774997896964, 2.2360679774997896964, 2.2360679774997896964, 1.0}};

constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }

// v moved between lanes by the DPP control CTRL (rows in ROW_MASK only;
// lanes without a source, or in other rows, get 0).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes, returned in every lane (a scalar register).  All
// lanes must be active.  DPP steps: within quads (lane ^ 1, lane ^ 2), row
// shifts by 4 and 8 (lane 15 of each row holds the row's sum), broadcast of
// lane 15 into the next row and of lane 31 into rows 2 and 3; lane 63 then
// holds the total.  A fixed order, so the result is deterministic.
__device__ __forceinline__ double wave_sum_all_active(double v)
{
    v += dpp_f64<0xb1, 0xf>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4e, 0xf>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x114, 0xf>(v);   // row_shr:4
    v += dpp_f64<0x118, 0xf>(v);   // row_shr:8
    v += dpp_f64<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, WAVE - 1);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), WAVE - 1);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ void atomic_add(QF_AS_GLOBAL double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// FP64 add into workgroup-shared memory (executed by the LDS unit).
__device__ __forceinline__ void lds_add(QF_AS_LDS double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------------------
// Contracted [e0|f0] for one (bra pair, ket pair): primitive loops with
// QUICK's primitive screen (dnmax*cutprim(ij))*cutprim(kl) > primLimit
// (eri.fh:270,288-289).  Primitive pairs are sorted by cutprim descending,
// so a failing bra primitive ends the inner loop and a failing first bra
// primitive ends the outer loop: the same set QUICK skips.  [pb, pe) and
// [qb, qe) restrict the bra and ket primitive pairs (a piece of a quartet
// spread over several tasks); the screen and its breaks are unchanged, as
// the first bra primitive still sets the outer test.
//
// The ket primitive (different per lane) is the outer loop: its seven
// fields are vector-loaded once per ket primitive (prefetching the next one
// cost 23 VGPRs and was slower); the bra primitives (the same for the whole
// wave) are the inner loop and come through scalar loads.
// ---------------------------------------------------------------------------
// grid: the FP32 Boys grid table (kBoysGrid or a copy in LDS), read by
// the FP32 instantiation only.
template <class C, typename R, bool BRA_UNIFORM, typename Tab, typename GT = const float*>
__host__ __device__ __forceinline__ void contract(const PairBlock& bb, int i, const PairInfo& bi, const PairBlock& kb,
                                         int j, const PairInfo& ki, double dn, double primlim, bool screen,
                                         int maxm, R (&acc)[C::NACC], Tab boys_tab, int pb, int pe, int qb,
                                         int qe, GT grid = kBoysGrid)
{
    const auto bp = BRA_UNIFORM ? (const double*)nullptr : bb.prim;
    const auto bpc = cptr(bb.prim);           // wave-uniform bra data
    const auto kp = gptr(kb.prim);            // per-lane ket data
    const int bn = bb.n, bK = bb.K, kn = kb.n, kK = kb.K;
    // Field f of primitive q of pair x lives at prim[(f K + q) n + x].  The
    // loads add a 32-bit byte offset (q n + x) * 8, shared by the fields, to
    // a wave-uniform field base: one address add per load instead of a
    // sign-extended 64-bit index computed per lane and field.
    const size_t bfs = (size_t)bK * bn, kfs = (size_t)kK * kn;
#define BP(f, q) (BRA_UNIFORM ? *(const QF_AS_CONST double*)((const QF_AS_CONST char*)(bpc + (f) * bfs) + \
                                                              (uint32_t)((q) * bn + i) * 8u)            \
                              : bp[(f) * bfs + (size_t)(q) * bn + i])
#define KP(f, off) (*(const QF_AS_GLOBAL double*)((const QF_AS_GLOBAL char*)(kp + (f) * kfs) + (off)))
    const double c0 = dn * BP(PF_CUT, 0);     // largest bra cutprim
    for (int jp = qb; jp < qe; ++jp) {
        const uint32_t ko = (uint32_t)(jp * kn + j) * 8u;
        const double cutk = KP(PF_CUT, ko);
        if (screen && !(c0 * cutk > primlim)) break;
        const double zk = KP(PF_ZETA, ko);
        const double ozk = KP(PF_OOZ2, ko);
        const double Kk = KP(PF_KC, ko);
        const double Qk[3] = {KP(PF_PX, ko), KP(PF_PY, ko), KP(PF_PZ, ko)};
        const double QC[3] = {Qk[0] - ki.A[0], Qk[1] - ki.A[1], Qk[2] - ki.A[2]};
        for (int ip0 = pb; ip0 < pe; ++ip0) {
            // the active lanes agree on ip: scalar loads of the bra data
            const int ip = BRA_UNIFORM ? wave_uniform(ip0) : ip0;
            if (screen && !((dn * BP(PF_CUT, ip)) * cutk > primlim)) break;
            const double zb = BP(PF_ZETA, ip);
            const double ozb = BP(PF_OOZ2, ip);
            const double Pb[3] = {BP(PF_PX, ip), BP(PF_PY, ip), BP(PF_PZ, ip)};
            const double Kb = BP(PF_KC, ip);
            const double PQ[3] = {Pb[0] - Qk[0], Pb[1] - Qk[1], Pb[2] - Qk[2]};
            R Fr[C::L + 1];
            VrrPar<R> p;
            if constexpr (sizeof(R) == 4) {
                // FP32 quartet: exponents, Boys function and VRR parameters in
                // FP32; only the coordinate differences are formed in FP64
                const float zbf = float(zb), zkf = float(zk);
                const float rs = qf_rsqrtf(zbf + zkf);
                const float inv = rs * rs;
                const float PQf[3] = {float(PQ[0]), float(PQ[1]), float(PQ[2])};
                const float rpq = fmaf(PQf[0], PQf[0], fmaf(PQf[1], PQf[1], PQf[2] * PQf[2]));
                float F[C::L + 1];
                boys_f32<C::L>(rpq * (zbf * zkf * inv), F, grid);
                const float pref = float(Kb * Kk) * rs;
#pragma unroll
                for (int m = 0; m <= C::L; ++m) Fr[m] = F[m] * pref;
                const float wp = zkf * inv, wq = zbf * inv;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    p.PA[d] = float(Pb[d] - bi.A[d]);
                    p.WP[d] = -wp * PQf[d];
                    p.QC[d] = float(QC[d]);
                    p.WQ[d] = wq * PQf[d];
                }
                p.oo2z = float(ozb);
                p.oo2n = float(ozk);
                p.oo2zn = 0.5f * inv;
                p.rz = wp;
                p.rn = wq;
            } else {
                const double rs = qf_rsqrt(zb + zk);       // 1/sqrt(zeta+eta)
                const double inv = rs * rs;
                const double rpq = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
                const double T = rpq * (zb * zk * inv);
                double F[C::L + 1];
                boys_quick<C::L>(T, maxm, F, boys_tab);
                const double pref = Kb * Kk * rs;
#pragma unroll
                for (int m = 0; m <= C::L; ++m) Fr[m] = R(F[m] * pref);
                const double wp = zk * inv, wq = zb * inv;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    p.PA[d] = R(Pb[d] - bi.A[d]);
                    p.WP[d] = R(-wp * PQ[d]);
                    p.QC[d] = R(QC[d]);
                    p.WQ[d] = R(wq * PQ[d]);
                }
                p.oo2z = R(ozb);
                p.oo2n = R(ozk);
                p.oo2zn = R(0.5 * inv);
                p.rz = R(wp);
                p.rn = R(wq);
            }
            C::template vrr<R>(Fr, p, acc);
        }
    }
#undef BP
#undef KP
}

// ---------------------------------------------------------------------------
// One task of the Fock build.  The 64 lanes hold up to 64 kets (c,d) of one
// ket block; the wave loops over bras (a,b_k) of one bra block that all
// share the first Q-shell a.  Screening per quartet as getCshellEri
// (eri.fh:173-189).  Digestion of the canonical quartet with degeneracy
// weight s = f/4, f = (a!=b ? 2:1)(c!=d ? 2:1)(ab!=cd ? 2:1):
//   Gt_ab += s D_cd I, Gt_cd += s D_ab I,
//   Gt_ac -= s/4 hyb D_bd I, Gt_ad -= .. D_bc I, Gt_bc -= .. D_ad I, Gt_bd -= .. D_ac I,
// and G = Gt + Gt^T afterwards.  Because a is fixed for the wave, J_cd, K_ac
// and K_ad stay in registers across the bra loop and J_ab is summed over
// the wave per bra.
//
// TILE = false: everything goes straight to G with FP64 atomics (K_bc and
// K_bd per quartet).  TILE = true: J_ab and every exchange term go to the
// workgroup's LDS tile instead (rows: functions of a, then nb functions per
// bra of the chunk; columns: the ket block's compact columns), and only
// J_cd is added to G here.  FP64 atomics on HBM are executed at the memory
// side at a fixed request rate, so the tile turns one atomic per quartet
// into one per tile element per workgroup.
// ---------------------------------------------------------------------------
struct TileRef {
    QF_AS_LDS double* k;     // nrow x ncol exchange tile
    QF_AS_LDS double* k2;    // the same for the second exchange density (open-shell pass)
    QF_AS_LDS double* jab;   // nbra x na x nb Coulomb terms of the bras
    int ncol, chunk_kb;
    const QF_AS_LDS double* boys;   // Boys fit table (kBoysFit) in LDS
    const QF_AS_LDS float* grid;    // FP32 Boys grid table (kBoysGrid) in LDS (FP32 kernels)
};

// OS = true (the fused open-shell pass): every integral is digested a second
// time into exchange terms of md.P2 that go to G2 (or the tile k2).  LEAN
// (tile kernels): the second exchange keeps no per-task state in registers;
// its a-row density blocks are re-read and its a-row terms added to the tile
// every quartet, so the contraction runs at the closed-shell register count.
// Classes with C::LEAN_K treat the first exchange the same way in their FP64
// tile kernels (the a-row blocks of P and terms of K_ac, K_ad).
template <class C, typename R, bool TILE, bool OS, bool LEAN = false>
__device__ __forceinline__ void fock_task(const Task& t, const PairBlock& bb, const PairBlock& kb,
                                          const MolDesc& md, QF_AS_GLOBAL double* G, QF_AS_GLOBAL double* G2,
                                          const TileRef& tr)
{
    constexpr int na = ncart(C::La), nb = ncart(C::Lb), nc = ncart(C::Lc), nd = ncart(C::Ld);
    constexpr bool kFP32 = sizeof(R) == 4;
    const int lane = threadIdx.x % WAVE;
    const bool in_range = lane < t.ket_count;
    const int j = in_range ? t.ket_start + lane : t.ket_start;
    const PairInfo ki = gptr(kb.info)[j];
    const int n = md.nbasis, ns = md.nshell;
    const auto cmc = cptr(md.cutmat);
    const auto cm = gptr(md.cutmat);
    const auto PJc = cptr(md.PJ);
    const auto PJ = gptr(md.PJ);
    const auto P = gptr(md.P);
    const auto bperm = cptr(bb.bperm);

    // the shared first Q-shell a of every bra of this task
    const PairInfo b0 = cptr(bb.info)[bperm[t.kb]];
    const int sa = b0.sa, fa = b0.fa;
    const double cm_cd = 4.0 * cm[ki.sa * ns + ki.sb];
    const double cm_ac = cm[sa * ns + ki.sa];
    const double cm_ad = cm[sa * ns + ki.sb];

    const auto P2 = gptr(md.P2);
    constexpr bool kLean = OS && TILE && LEAN;
    // the same for the first exchange, per class (C::LEAN_K, FP64 tile kernels)
    constexpr bool kLeanK = TILE && C::LEAN_K && !kFP32;
    constexpr int t1 = kLeanK ? 0 : 1;
    constexpr int os2 = (OS && !kLean) ? 1 : 0;   // sizes of the per-task second-density arrays (1 when unused)
    constexpr int osq = OS ? 1 : 0;               // the same for the per-quartet arrays
    double Dcd[nc * nd], Dac[t1 * na * nc + 1 - t1], Dad[t1 * na * nd + 1 - t1];
    double Dac2[os2 * na * nc + 1 - os2], Dad2[os2 * na * nd + 1 - os2];
#pragma unroll
    for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d) Dcd[c * nd + d] = PJ[(ki.fa + c) * n + ki.fb + d];
#pragma unroll
    for (int a = 0; a < na; ++a) {
        if constexpr (!kLeanK) {
#pragma unroll
            for (int c = 0; c < nc; ++c) Dac[a * nc + c] = P[(fa + a) * n + ki.fa + c];
#pragma unroll
            for (int d = 0; d < nd; ++d) Dad[a * nd + d] = P[(fa + a) * n + ki.fb + d];
        }
        if constexpr (OS && !kLean) {
#pragma unroll
            for (int c = 0; c < nc; ++c) Dac2[a * nc + c] = P2[(fa + a) * n + ki.fa + c];
#pragma unroll
            for (int d = 0; d < nd; ++d) Dad2[a * nd + d] = P2[(fa + a) * n + ki.fb + d];
        }
    }
    double Jcd[nc * nd], Kac[t1 * na * nc + 1 - t1], Kad[t1 * na * nd + 1 - t1];
    double Kac2[os2 * na * nc + 1 - os2], Kad2[os2 * na * nd + 1 - os2];
#pragma unroll
    for (int k = 0; k < nc * nd; ++k) Jcd[k] = 0.0;
    if constexpr (!kLeanK) {
#pragma unroll
        for (int k = 0; k < na * nc; ++k) Kac[k] = 0.0;
#pragma unroll
        for (int k = 0; k < na * nd; ++k) Kad[k] = 0.0;
    }
    if constexpr (OS && !kLean) {
#pragma unroll
        for (int k = 0; k < na * nc; ++k) Kac2[k] = 0.0;
#pragma unroll
        for (int k = 0; k < na * nd; ++k) Kad2[k] = 0.0;
    }
    bool touched = false;
    const double kscale = -0.25 * md.hyb;

    for (int kk0 = t.kb; kk0 < t.ke; ++kk0) {
        const int kk = wave_uniform(kk0);
        const int i = bperm[kk];
        const PairInfo bi = cptr(bb.info)[i];
        const double ytest = bi.ycut * ki.ycut;
        double dn = fmax(4.0 * cmc[bi.sa * ns + bi.sb], cm_cd);
        dn = fmax(dn, cm_ad);
        dn = fmax(dn, cm_ac);
        dn = fmax(dn, cm[bi.sb * ns + ki.sa]);
        dn = fmax(dn, cm[bi.sb * ns + ki.sb]);
        bool valid = in_range && (ytest > md.cutoff) && (ytest * dn > md.cutoff) && (!t.same_block || j <= i);
        valid = valid && (kFP32 ? (ytest < md.split) : (t.all_fp64 || ytest >= md.split));
        if (!__any(valid)) continue;
        touched = touched || valid;

        R acc[C::NACC];
#pragma unroll
        for (int k = 0; k < C::NACC; ++k) acc[k] = R(0);
        if (valid)
            contract<C, R, true>(bb, i, bi, kb, j, ki, dn, md.primlim, true, bi.qfs + ki.qfs, acc, tr.boys, t.pb, t.pe,
                                 t.qb, t.qe, tr.grid);
        R I[C::NI];
        {
            const R AB[3] = {R(bi.AB[0]), R(bi.AB[1]), R(bi.AB[2])};
            const R CD[3] = {R(ki.AB[0]), R(ki.AB[1]), R(ki.AB[2])};
            C::template hrr<R>(acc, AB, CD, I);
        }
        const double s = valid ? 0.25 * (bi.qa != bi.qb ? 2.0 : 1.0) * (ki.qa != ki.qb ? 2.0 : 1.0) *
                                     (bi.id != ki.id ? 2.0 : 1.0)
                               : 0.0;
        double Dab[na * nb], Dbc[nb * nc], Dbd[nb * nd];
#pragma unroll
        for (int a = 0; a < na; ++a)
#pragma unroll
            for (int b = 0; b < nb; ++b) Dab[a * nb + b] = PJc[(bi.fa + a) * n + bi.fb + b];
#pragma unroll
        for (int b = 0; b < nb; ++b) {
#pragma unroll
            for (int c = 0; c < nc; ++c) Dbc[b * nc + c] = P[(bi.fb + b) * n + ki.fa + c];
#pragma unroll
            for (int d = 0; d < nd; ++d) Dbd[b * nd + d] = P[(bi.fb + b) * n + ki.fb + d];
        }
        double Jab[na * nb], Kbc[nb * nc], Kbd[nb * nd];
        double Dbc2[osq * nb * nc + 1 - osq], Dbd2[osq * nb * nd + 1 - osq];
        double Kbc2[osq * nb * nc + 1 - osq], Kbd2[osq * nb * nd + 1 - osq];
        constexpr int lq = kLean ? 1 : 0;
        double Dac2q[lq * na * nc + 1 - lq], Dad2q[lq * na * nd + 1 - lq];
        constexpr int lk = kLeanK ? 1 : 0;
        double Dacq[lk * na * nc + 1 - lk], Dadq[lk * na * nd + 1 - lk];
        double Kacq[lk * na * nc + 1 - lk], Kadq[lk * na * nd + 1 - lk];
        if constexpr (kLeanK) {
            // the a-row blocks of P again for this quartet, from offsets the
            // compiler cannot hoist out of the bra loop (see the P2 blocks below)
            int oc = fa * n + ki.fa, od = fa * n + ki.fb;
            asm volatile("" : "+v"(oc), "+v"(od));
#pragma unroll
            for (int a = 0; a < na; ++a) {
#pragma unroll
                for (int c = 0; c < nc; ++c) Dacq[a * nc + c] = P[oc + a * n + c];
#pragma unroll
                for (int d = 0; d < nd; ++d) Dadq[a * nd + d] = P[od + a * n + d];
            }
#pragma unroll
            for (int k = 0; k < na * nc; ++k) Kacq[k] = 0.0;
#pragma unroll
            for (int k = 0; k < na * nd; ++k) Kadq[k] = 0.0;
        }
        double Kac2q[lq * na * nc + 1 - lq], Kad2q[lq * na * nd + 1 - lq];
#pragma unroll
        for (int k = 0; k < na * nb; ++k) Jab[k] = 0.0;
#pragma unroll
        for (int k = 0; k < nb * nc; ++k) Kbc[k] = 0.0;
#pragma unroll
        for (int k = 0; k < nb * nd; ++k) Kbd[k] = 0.0;
        if constexpr (OS) {
#pragma unroll
            for (int b = 0; b < nb; ++b) {
#pragma unroll
                for (int c = 0; c < nc; ++c) Dbc2[b * nc + c] = P2[(bi.fb + b) * n + ki.fa + c];
#pragma unroll
                for (int d = 0; d < nd; ++d) Dbd2[b * nd + d] = P2[(bi.fb + b) * n + ki.fb + d];
            }
#pragma unroll
            for (int k = 0; k < nb * nc; ++k) Kbc2[k] = 0.0;
#pragma unroll
            for (int k = 0; k < nb * nd; ++k) Kbd2[k] = 0.0;
        }
        if constexpr (kLean) {
            // opaque offsets: the blocks are the same every quartet, and a
            // hoisted load would hold them in registers across the loop again
            int oc = fa * n + ki.fa, od = fa * n + ki.fb;
            asm volatile("" : "+v"(oc), "+v"(od));
#pragma unroll
            for (int a = 0; a < na; ++a) {
#pragma unroll
                for (int c = 0; c < nc; ++c) Dac2q[a * nc + c] = P2[oc + a * n + c];
#pragma unroll
                for (int d = 0; d < nd; ++d) Dad2q[a * nd + d] = P2[od + a * n + d];
            }
#pragma unroll
            for (int k = 0; k < na * nc; ++k) Kac2q[k] = 0.0;
#pragma unroll
            for (int k = 0; k < na * nd; ++k) Kad2q[k] = 0.0;
        }
#pragma unroll
        for (int a = 0; a < na; ++a)
#pragma unroll
            for (int b = 0; b < nb; ++b)
#pragma unroll
                for (int c = 0; c < nc; ++c)
#pragma unroll
                    for (int d = 0; d < nd; ++d) {
                        const double cons = kCons[C::La][a] * kCons[C::Lb][b] * kCons[C::Lc][c] * kCons[C::Ld][d];
                        const double v = double(I[((a * nb + b) * nc + c) * nd + d]) * (cons * s);
                        Jab[a * nb + b] += Dcd[c * nd + d] * v;
                        Jcd[c * nd + d] += Dab[a * nb + b] * v;
                        if constexpr (kLeanK) {
                            Kacq[a * nc + c] += Dbd[b * nd + d] * v;
                            Kadq[a * nd + d] += Dbc[b * nc + c] * v;
                            Kbc[b * nc + c] += Dadq[a * nd + d] * v;
                            Kbd[b * nd + d] += Dacq[a * nc + c] * v;
                        } else {
                            Kac[a * nc + c] += Dbd[b * nd + d] * v;
                            Kad[a * nd + d] += Dbc[b * nc + c] * v;
                            Kbc[b * nc + c] += Dad[a * nd + d] * v;
                            Kbd[b * nd + d] += Dac[a * nc + c] * v;
                        }
                        if constexpr (kLean) {
                            Kac2q[a * nc + c] += Dbd2[b * nd + d] * v;
                            Kad2q[a * nd + d] += Dbc2[b * nc + c] * v;
                            Kbc2[b * nc + c] += Dad2q[a * nd + d] * v;
                            Kbd2[b * nd + d] += Dac2q[a * nc + c] * v;
                        } else if constexpr (OS) {
                            Kac2[a * nc + c] += Dbd2[b * nd + d] * v;
                            Kad2[a * nd + d] += Dbc2[b * nc + c] * v;
                            Kbc2[b * nc + c] += Dad2[a * nd + d] * v;
                            Kbd2[b * nd + d] += Dac2[a * nc + c] * v;
                        }
                    }
        // bra block: summed over the wave
#pragma unroll
        for (int a = 0; a < na; ++a)
#pragma unroll
            for (int b = 0; b < nb; ++b) {
                const double v = wave_sum_all_active(Jab[a * nb + b]);
                if (lane == 0 && v != 0.0) {
                    if (TILE)
                        lds_add(&tr.jab[((kk - tr.chunk_kb) * na + a) * nb + b], v);
                    else
                        atomic_add(&G[(bi.fa + a) * n + bi.fb + b], v);
                }
            }
        if (valid) {
            if (TILE) {
                const int r0 = na + (kk - tr.chunk_kb) * nb;
#pragma unroll
                for (int b = 0; b < nb; ++b) {
                    QF_AS_LDS double* row = tr.k + (r0 + b) * tr.ncol;
#pragma unroll
                    for (int c = 0; c < nc; ++c) lds_add(&row[ki.cola + c], Kbc[b * nc + c]);
#pragma unroll
                    for (int d = 0; d < nd; ++d) lds_add(&row[ki.colb + d], Kbd[b * nd + d]);
                    if constexpr (OS) {
                        if (tr.k2) {
                            QF_AS_LDS double* row2 = tr.k2 + (r0 + b) * tr.ncol;
#pragma unroll
                            for (int c = 0; c < nc; ++c) lds_add(&row2[ki.cola + c], Kbc2[b * nc + c]);
#pragma unroll
                            for (int d = 0; d < nd; ++d) lds_add(&row2[ki.colb + d], Kbd2[b * nd + d]);
                        } else {
#pragma unroll
                            for (int c = 0; c < nc; ++c)
                                atomic_add(&G2[(bi.fb + b) * n + ki.fa + c], kscale * Kbc2[b * nc + c]);
#pragma unroll
                            for (int d = 0; d < nd; ++d)
                                atomic_add(&G2[(bi.fb + b) * n + ki.fb + d], kscale * Kbd2[b * nd + d]);
                        }
                    }
                }
                if constexpr (kLeanK) {
#pragma unroll
                    for (int a = 0; a < na; ++a) {
                        QF_AS_LDS double* row = tr.k + a * tr.ncol;
#pragma unroll
                        for (int c = 0; c < nc; ++c) lds_add(&row[ki.cola + c], Kacq[a * nc + c]);
#pragma unroll
                        for (int d = 0; d < nd; ++d) lds_add(&row[ki.colb + d], Kadq[a * nd + d]);
                    }
                }
                if constexpr (kLean) {
#pragma unroll
                    for (int a = 0; a < na; ++a) {
                        if (tr.k2) {
                            QF_AS_LDS double* row2 = tr.k2 + a * tr.ncol;
#pragma unroll
                            for (int c = 0; c < nc; ++c) lds_add(&row2[ki.cola + c], Kac2q[a * nc + c]);
#pragma unroll
                            for (int d = 0; d < nd; ++d) lds_add(&row2[ki.colb + d], Kad2q[a * nd + d]);
                        } else {
#pragma unroll
                            for (int c = 0; c < nc; ++c)
                                atomic_add(&G2[(fa + a) * n + ki.fa + c], kscale * Kac2q[a * nc + c]);
#pragma unroll
                            for (int d = 0; d < nd; ++d)
                                atomic_add(&G2[(fa + a) * n + ki.fb + d], kscale * Kad2q[a * nd + d]);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int b = 0; b < nb; ++b) {
#pragma unroll
                    for (int c = 0; c < nc; ++c)
                        atomic_add(&G[(bi.fb + b) * n + ki.fa + c], kscale * Kbc[b * nc + c]);
#pragma unroll
                    for (int d = 0; d < nd; ++d)
                        atomic_add(&G[(bi.fb + b) * n + ki.fb + d], kscale * Kbd[b * nd + d]);
                    if constexpr (OS) {
#pragma unroll
                        for (int c = 0; c < nc; ++c)
                            atomic_add(&G2[(bi.fb + b) * n + ki.fa + c], kscale * Kbc2[b * nc + c]);
#pragma unroll
                        for (int d = 0; d < nd; ++d)
                            atomic_add(&G2[(bi.fb + b) * n + ki.fb + d], kscale * Kbd2[b * nd + d]);
                    }
                }
            }
        }
    }
    if (!touched) return;
#pragma unroll
    for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d) atomic_add(&G[(ki.fa + c) * n + ki.fb + d], Jcd[c * nd + d]);
    if (TILE) {
#pragma unroll
        for (int a = 0; a < na; ++a) {
            QF_AS_LDS double* row = tr.k + a * tr.ncol;
            if constexpr (!kLeanK) {
#pragma unroll
                for (int c = 0; c < nc; ++c) lds_add(&row[ki.cola + c], Kac[a * nc + c]);
#pragma unroll
                for (int d = 0; d < nd; ++d) lds_add(&row[ki.colb + d], Kad[a * nd + d]);
            }
            if constexpr (OS && !kLean) {
                if (tr.k2) {
                    QF_AS_LDS double* row2 = tr.k2 + a * tr.ncol;
#pragma unroll
                    for (int c = 0; c < nc; ++c) lds_add(&row2[ki.cola + c], Kac2[a * nc + c]);
#pragma unroll
                    for (int d = 0; d < nd; ++d) lds_add(&row2[ki.colb + d], Kad2[a * nd + d]);
                } else {
#pragma unroll
                    for (int c = 0; c < nc; ++c)
                        atomic_add(&G2[(fa + a) * n + ki.fa + c], kscale * Kac2[a * nc + c]);
#pragma unroll
                    for (int d = 0; d < nd; ++d)
                        atomic_add(&G2[(fa + a) * n + ki.fb + d], kscale * Kad2[a * nd + d]);
                }
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < na; ++a) {
#pragma unroll
            for (int c = 0; c < nc; ++c) atomic_add(&G[(fa + a) * n + ki.fa + c], kscale * Kac[a * nc + c]);
#pragma unroll
            for (int d = 0; d < nd; ++d) atomic_add(&G[(fa + a) * n + ki.fb + d], kscale * Kad[a * nd + d]);
            if constexpr (OS) {
#pragma unroll
                for (int c = 0; c < nc; ++c) atomic_add(&G2[(fa + a) * n + ki.fa + c], kscale * Kac2[a * nc + c]);
#pragma unroll
                for (int d = 0; d < nd; ++d) atomic_add(&G2[(fa + a) * n + ki.fb + d], kscale * Kad2[a * nd + d]);
            }
        }
    }
}

// One wave per task, every term straight to G (replica blockIdx % replicas).
// The grid may be smaller than the task count: each workgroup then strides
// over the tasks (a persistent launch).
template <class C, typename R, bool OS = false>
__global__ __attribute__((amdgpu_flat_work_group_size(WAVE, WAVE), amdgpu_waves_per_eu(OS ? C::MIN_WAVES_OS : C::MIN_WAVES)))
void fock_kernel(const Task* __restrict__ tasks, int ntasks, const PairBlock* __restrict__ blocks,
                 const MolDesc* __restrict__ mols)
{
    QF_BOYS_LDS_DECL;
    QF_BOYS_LDS_FILL();
    TileRef tr{};
    tr.boys = QF_BOYS_LDS_PTR;
    for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const Task t = cptr(tasks)[ti];
        const PairBlock bb = cptr(blocks)[t.bra_block];
        const PairBlock kb = cptr(blocks)[t.ket_block];
        const MolDesc md = cptr(mols)[bb.mol];
        // spread the atomics of concurrent waves over several copies of G
        const size_t roff = (size_t)(blockIdx.x % md.replicas) * md.nbasis * md.nbasis;
        const auto G = gptr_rw(md.G + roff);
        const auto G2 = gptr_rw(OS ? md.G2 + roff : md.G);
        fock_task<C, R, false, OS>(t, bb, kb, md, G, G2, tr);
    }
}

// One workgroup per TileGroup (striding over the groups when the grid is
// smaller): its waves take the group's tasks in turn and accumulate J_ab
// and the exchange terms in LDS; after a barrier the workgroup adds the
// non-zero tile elements to G.  Dynamic LDS:
// (na + nbra*nb) * ncol * (k2 ? 2 : 1) + nbra*na*nb doubles.  OS: the
// second exchange terms go to a second LDS tile (k2 = 1) or straight to
// G2 (k2 = 0, where two tiles do not fit).
template <class C, typename R, bool OS = false, bool LEAN = false>
__global__ __attribute__((amdgpu_flat_work_group_size(WAVE, 4 * WAVE),
                          amdgpu_waves_per_eu((OS && !LEAN) ? C::MIN_WAVES_OS : C::MIN_WAVES)))
void fock_tile_kernel(const TileGroup* __restrict__ groups, int ngroups, const Task* __restrict__ tasks,
                      const PairBlock* __restrict__ blocks, const MolDesc* __restrict__ mols, int k2)
{
    constexpr int na = ncart(C::La), nb = ncart(C::Lb);
    constexpr bool kF32 = sizeof(R) == 4;
    extern __shared__ double qf_lds[];
    // FP32 kernels: the FP32 Boys grid in LDS (6 KB; global-memory gathers
    // per primitive stalled the loop)
    __shared__ float qf_boys_grid[kF32 ? kBoysGridN * kBoysGridM : 1];
    if constexpr (kF32)
        for (int x = threadIdx.x; x < kBoysGridN * kBoysGridM; x += blockDim.x) qf_boys_grid[x] = kBoysGrid[x];
    QF_BOYS_LDS_DECL;
    QF_BOYS_LDS_FILL();
    QF_AS_LDS double* lds = (QF_AS_LDS double*)qf_lds;
    const int nw = blockDim.x / WAVE;
    for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const TileGroup g = cptr(groups)[gi];
        const Task t0 = cptr(tasks)[g.task_begin];
        const PairBlock bb = cptr(blocks)[t0.bra_block];
        const PairBlock kb = cptr(blocks)[t0.ket_block];
        const MolDesc md = cptr(mols)[bb.mol];
        const int n = md.nbasis;
        const int nrow = na + g.nbra * nb, ncol = kb.ncol;
        const int ntile = nrow * ncol, njab = g.nbra * na * nb;
        const bool two = OS && k2;
        const int nk = two ? 2 * ntile : ntile;
        for (int x = threadIdx.x; x < nk + njab; x += blockDim.x) lds[x] = 0.0;
        __syncthreads();

        const size_t roff = (size_t)(blockIdx.x % md.replicas) * n * n;
        const auto G = gptr_rw(md.G + roff);
        const auto G2 = gptr_rw(OS ? md.G2 + roff : md.G);
        const TileRef tr{lds, two ? lds + ntile : nullptr, lds + nk, ncol, g.chunk_kb, QF_BOYS_LDS_PTR,
                         (const QF_AS_LDS float*)qf_boys_grid};
        // the wave index through readfirstlane: the compiler cannot prove
        // threadIdx.x / 64 wave-uniform, and would otherwise keep the task,
        // its loop bounds and the bra primitive data in vector registers
        const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
        for (int ti = g.task_begin + wave; ti < g.task_end; ti += nw) {
            const Task t = cptr(tasks)[wave_uniform(ti)];
            fock_task<C, R, true, OS, LEAN>(t, bb, kb, md, G, G2, tr);
        }
        __syncthreads();

        // flush: row r -> function of a (r < na) or of b_k; column q -> colf[q]
        const double kscale = -0.25 * md.hyb;
        const auto bperm = cptr(bb.bperm);
        const auto colf = gptr(kb.colf);
        const int fa = cptr(bb.info)[bperm[g.chunk_kb]].fa;
        for (int x = threadIdx.x; x < nk; x += blockDim.x) {
            const double v = lds[x];
            if (v == 0.0) continue;
            const int xt = x < ntile ? x : x - ntile;
            const int r = xt / ncol, q = xt - r * ncol;
            int fr;
            if (r < na) {
                fr = fa + r;
            } else {
                const int k = (r - na) / nb;
                fr = gptr(bb.info)[bperm[g.chunk_kb + k]].fb + (r - na - k * nb);
            }
            atomic_add(&(x < ntile ? G : G2)[fr * n + colf[q]], kscale * v);
        }
        for (int x = threadIdx.x; x < njab; x += blockDim.x) {
            const double v = lds[nk + x];
            if (v == 0.0) continue;
            const int k = x / (na * nb), ab = x - k * (na * nb);
            const PairInfo bi = gptr(bb.info)[bperm[g.chunk_kb + k]];
            atomic_add(&G[(bi.fa + ab / nb) * n + bi.fb + ab % nb], v);
        }
        __syncthreads();   // the next group reuses the tile
    }
}

// ---------------------------------------------------------------------------
// Schwarz kernel for a diagonal class (La Lb | La Lb): one lane per pair,
// out[j] = max over components |(ab|ab)| with no screening
// (shellcutoff / classcutoff, quick_cutoff_module.f90:91-424).  Used with
// contracted pairs (Ycutoff) and with single-primitive pairs (cutprim).
// ---------------------------------------------------------------------------
template <class C>
__global__ __launch_bounds__(WAVE) void schwarz_kernel(PairBlock blk, double* __restrict__ out)
{
    constexpr int na = ncart(C::La), nb = ncart(C::Lb);
    QF_BOYS_LDS_DECL;
    QF_BOYS_LDS_FILL();
    const int j = blockIdx.x * WAVE + threadIdx.x;
    if (j >= blk.n) return;
    const PairInfo pi = blk.info[j];
    double acc[C::NACC];
#pragma unroll
    for (int k = 0; k < C::NACC; ++k) acc[k] = 0.0;
    contract<C, double, false>(blk, j, pi, blk, j, pi, 0.0, 0.0, false, 2 * pi.qfs, acc, QF_BOYS_LDS_PTR, 0, blk.K, 0,
                               blk.K);
    double I[C::NI];
    const double AB[3] = {pi.AB[0], pi.AB[1], pi.AB[2]};
    C::template hrr<double>(acc, AB, AB, I);
    double m = 0.0;
#pragma unroll
    for (int a = 0; a < na; ++a)
#pragma unroll
        for (int b = 0; b < nb; ++b) {
            const double cons = kCons[C::La][a] * kCons[C::Lb][b];
            m = fmax(m, fabs(I[((a * nb + b) * na + a) * nb + b] * cons * cons));
        }
    out[j] = m;
}

}  // namespace qf
