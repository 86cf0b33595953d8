// One-electron matrices S, T, V for the SCF harness around the Fock build
// (QUICK: quick_overlap_module.f90:172-363, quick_oei_module.f90:69-257,
// attrashell.fh).  One thread per basis-function pair (s through f),
// McMurchie-Davidson
// Hermite expansion; the nuclear-attraction Boys function is QUICK's FmT
// called with Maxm = Qfinal(I) + Qfinal(J) as attrashell.fh:46-48,89.
// Runs once per geometry, so it favours simplicity over speed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/quickfock.h"
#include "qf_device.h"

namespace {

struct OeiFn {
    int shell, L, lx, ly, lz, pstart, nprim, lmaxs;
    double cons;
    double X, Y, Z;
};

// Hermite coefficients E^{i,j}_t (i <= 3, j <= 5, t <= 8) in one direction,
// without the exp(-mu X_AB^2) factor.
constexpr int kEI = 4, kEJ = 6, kET = 9;
__device__ void hermite_table(int la, int lb, double XPA, double XPB, double oo2p, double (&E)[kEI][kEJ][kET])
{
    for (int i = 0; i < kEI; ++i)
        for (int j = 0; j < kEJ; ++j)
            for (int t = 0; t < kET; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (i > 0) {
                    v = XPA * E[i - 1][j][t] + (t + 1 < kET ? (t + 1) * E[i - 1][j][t + 1] : 0.0);
                    if (t > 0) v += oo2p * E[i - 1][j][t - 1];
                } else {
                    v = XPB * E[i][j - 1][t] + (t + 1 < kET ? (t + 1) * E[i][j - 1][t + 1] : 0.0);
                    if (t > 0) v += oo2p * E[i][j - 1][t - 1];
                }
                E[i][j][t] = v;
            }
        }
}

constexpr int kLV = 6;   // highest La + Lb of the nuclear attraction (f f)

template <int L>
__device__ void boys_copy(double X, int maxm, double (&G)[kLV + 1])
{
    double H[L + 1];
    qf::boys_quick<L>(X, maxm, H, qf::kBoysFit);
    for (int m = 0; m <= L; ++m) G[m] = H[m];
}

__device__ void boys_rt(double X, int maxm, int L, double (&F)[kLV + 1])
{
    // same algorithm as qf::boys_quick, with a run-time order (L <= 6)
    double G[kLV + 1] = {};
    switch (L) {
        case 0: boys_copy<0>(X, maxm, G); break;
        case 1: boys_copy<1>(X, maxm, G); break;
        case 2: boys_copy<2>(X, maxm, G); break;
        case 3: boys_copy<3>(X, maxm, G); break;
        case 4: boys_copy<4>(X, maxm, G); break;
        case 5: boys_copy<5>(X, maxm, G); break;
        default: boys_copy<6>(X, maxm, G); break;
    }
    for (int m = 0; m <= kLV; ++m) F[m] = m <= L ? G[m] : 0.0;
}

__global__ void k_oei(const OeiFn* __restrict__ fn, int n, const double* __restrict__ exps,
                      const double* __restrict__ coef, int natom, const double* __restrict__ xyz,
                      const double* __restrict__ charge, double* __restrict__ S, double* __restrict__ T,
                      double* __restrict__ V)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int mu = idx / n, nu = idx % n;
    const OeiFn a = fn[mu], b = fn[nu];
    const int la[3] = {a.lx, a.ly, a.lz}, lb[3] = {b.lx, b.ly, b.lz};
    const double A[3] = {a.X, a.Y, a.Z}, B[3] = {b.X, b.Y, b.Z};
    const int L = a.L + b.L;
    const int maxm = a.lmaxs + b.lmaxs;
    double AB2 = 0.0;
    for (int d = 0; d < 3; ++d) AB2 += (A[d] - B[d]) * (A[d] - B[d]);
    double s_sum = 0.0, t_sum = 0.0, v_sum = 0.0;
    for (int i = 0; i < a.nprim; ++i)
        for (int j = 0; j < b.nprim; ++j) {
            const double al = exps[a.pstart + i], be = exps[b.pstart + j];
            const double ca = coef[(a.pstart + i) * 4 + a.L] * a.cons;
            const double cb = coef[(b.pstart + j) * 4 + b.L] * b.cons;
            const double p = al + be, oo2p = 0.5 / p;
            const double Kab = exp(-al * be / p * AB2);
            double P[3], E[3][kEI][kEJ][kET];
            for (int d = 0; d < 3; ++d) {
                P[d] = (al * A[d] + be * B[d]) / p;
                hermite_table(la[d], lb[d] + 2, P[d] - A[d], P[d] - B[d], oo2p, E[d]);
            }
            const double pre = pow(3.14159265358979323846 / p, 1.5) * Kab;
            double s1[3], t1[3];
            for (int d = 0; d < 3; ++d) {
                const int ai = la[d], bj = lb[d];
                s1[d] = E[d][ai][bj][0];
                double tk = -2.0 * be * be * E[d][ai][bj + 2][0] + be * (2 * bj + 1) * E[d][ai][bj][0];
                if (bj >= 2) tk -= 0.5 * bj * (bj - 1) * E[d][ai][bj - 2][0];
                t1[d] = tk;
            }
            s_sum += ca * cb * pre * s1[0] * s1[1] * s1[2];
            t_sum += ca * cb * pre * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
            // nuclear attraction: -Z 2pi/p K_AB sum_tuv E^x_t E^y_u E^z_v R_tuv(P - C)
            for (int c = 0; c < natom; ++c) {
                const double PC[3] = {P[0] - xyz[3 * c], P[1] - xyz[3 * c + 1], P[2] - xyz[3 * c + 2]};
                const double Tb = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                double F[kLV + 1];
                boys_rt(Tb, maxm, L, F);
                // R^N_tuv level by level, N = L..0
                double R[kLV + 1][kLV + 1][kLV + 1], Rn[kLV + 1][kLV + 1][kLV + 1];
                for (int N = L; N >= 0; --N) {
                    double f = F[N];
                    for (int k = 0; k < N; ++k) f *= -2.0 * p;
                    for (int t = 0; t <= L - N; ++t)
                        for (int u = 0; u <= L - N - t; ++u)
                            for (int v = 0; v <= L - N - t - u; ++v) {
                                double r;
                                if (t + u + v == 0) {
                                    r = f;
                                } else if (t > 0) {
                                    r = PC[0] * R[t - 1][u][v] + (t > 1 ? (t - 1) * R[t - 2][u][v] : 0.0);
                                } else if (u > 0) {
                                    r = PC[1] * R[t][u - 1][v] + (u > 1 ? (u - 1) * R[t][u - 2][v] : 0.0);
                                } else {
                                    r = PC[2] * R[t][u][v - 1] + (v > 1 ? (v - 1) * R[t][u][v - 2] : 0.0);
                                }
                                Rn[t][u][v] = r;
                            }
                    for (int t = 0; t <= L - N; ++t)
                        for (int u = 0; u <= L - N - t; ++u)
                            for (int v = 0; v <= L - N - t - u; ++v) R[t][u][v] = Rn[t][u][v];
                }
                double sum = 0.0;
                for (int t = 0; t <= la[0] + lb[0]; ++t)
                    for (int u = 0; u <= la[1] + lb[1]; ++u)
                        for (int v = 0; v <= la[2] + lb[2]; ++v)
                            sum += E[0][la[0]][lb[0]][t] * E[1][la[1]][lb[1]][u] * E[2][la[2]][lb[2]][v] * R[t][u][v];
                v_sum -= charge[c] * ca * cb * 2.0 * 3.14159265358979323846 / p * Kab * sum;
            }
        }
    S[idx] = s_sum;
    T[idx] = t_sum;
    V[idx] = v_sum;
}

}  // namespace

// The engine's basis is reachable through a small accessor declared in engine.hip.
extern "C" int qf_internal_basis(qf_engine* e, int* natom, const double** xyz, int* nshell, const int** s_atom,
                                 const int** s_lmin, const int** s_lmax, const int** s_nprim, const int** s_pstart,
                                 const int** s_fstart, int* nprim, const double** exps, const double** coef,
                                 int* nbasis);

extern "C" int qf_one_electron_ext(qf_engine* e, int ncenter, const double* center_xyz, const double* charges,
                                   double* S, double* T, double* V)
{
    int natom, nshell, nprim, nbasis;
    const double *xyz, *exps, *coef;
    const int *s_atom, *s_lmin, *s_lmax, *s_nprim, *s_pstart, *s_fstart;
    int rc = qf_internal_basis(e, &natom, &xyz, &nshell, &s_atom, &s_lmin, &s_lmax, &s_nprim, &s_pstart, &s_fstart,
                               &nprim, &exps, &coef, &nbasis);
    if (rc) return rc;
    if (ncenter < 0 || (ncenter > 0 && (!center_xyz || !charges)) || !S || !T || !V) return QF_ERR_ARG;
    // QUICK's component order and cons per shell type (basis.f90:567-720)
    static const int comp[4][10][3] = {
        {{0, 0, 0}},
        {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
        {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}},
        {{3, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 3, 0}, {2, 0, 1}, {1, 1, 1}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {0, 0, 3}}};
    const double s3 = std::sqrt(3.0), s5 = std::sqrt(5.0), s15 = std::sqrt(15.0);
    const double cons[4][10] = {{1}, {1, 1, 1}, {1, s3, 1, s3, s3, 1}, {1, s5, s5, 1, s5, s15, s5, s5, s5, 1}};
    std::vector<OeiFn> fn(nbasis);
    for (int s = 0; s < nshell; ++s) {
        int f = s_fstart[s];
        for (int L = s_lmin[s]; L <= s_lmax[s]; ++L)
            for (int k = 0; k < (L + 1) * (L + 2) / 2; ++k, ++f) {
                const double* X = xyz + 3 * s_atom[s];
                fn[f] = OeiFn{s, L, comp[L][k][0], comp[L][k][1], comp[L][k][2], s_pstart[s], s_nprim[s], s_lmax[s],
                              cons[L][k], X[0], X[1], X[2]};
            }
    }
    const size_t n2 = (size_t)nbasis * nbasis;
    const int nc = std::max(ncenter, 1);
    OeiFn* dfn = nullptr;
    double *dex = nullptr, *dco = nullptr, *dxyz = nullptr, *dch = nullptr, *dS = nullptr;
    hipError_t err = hipMalloc(&dfn, sizeof(OeiFn) * nbasis);
    if (err == hipSuccess) err = hipMalloc(&dex, sizeof(double) * nprim);
    if (err == hipSuccess) err = hipMalloc(&dco, sizeof(double) * 4 * nprim);
    if (err == hipSuccess) err = hipMalloc(&dxyz, sizeof(double) * 3 * nc);
    if (err == hipSuccess) err = hipMalloc(&dch, sizeof(double) * nc);
    if (err == hipSuccess) err = hipMalloc(&dS, sizeof(double) * n2 * 3);
    if (err == hipSuccess) err = hipMemcpy(dfn, fn.data(), sizeof(OeiFn) * nbasis, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(dex, exps, sizeof(double) * nprim, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(dco, coef, sizeof(double) * 4 * nprim, hipMemcpyHostToDevice);
    if (err == hipSuccess && ncenter > 0) err = hipMemcpy(dxyz, center_xyz, sizeof(double) * 3 * ncenter, hipMemcpyHostToDevice);
    if (err == hipSuccess && ncenter > 0) err = hipMemcpy(dch, charges, sizeof(double) * ncenter, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        double* dT = dS + n2;
        double* dV = dT + n2;
        hipLaunchKernelGGL(k_oei, dim3((n2 + 127) / 128), dim3(128), 0, 0, dfn, nbasis, dex, dco, ncenter, dxyz, dch,
                           dS, dT, dV);
        err = hipGetLastError();
        if (err == hipSuccess) err = hipMemcpy(S, dS, sizeof(double) * n2, hipMemcpyDeviceToHost);
        if (err == hipSuccess) err = hipMemcpy(T, dT, sizeof(double) * n2, hipMemcpyDeviceToHost);
        if (err == hipSuccess) err = hipMemcpy(V, dV, sizeof(double) * n2, hipMemcpyDeviceToHost);
    }
    (void)hipFree(dfn);
    (void)hipFree(dex);
    (void)hipFree(dco);
    (void)hipFree(dxyz);
    (void)hipFree(dch);
    (void)hipFree(dS);
    return err == hipSuccess ? QF_OK : QF_ERR_HIP;
}

extern "C" int qf_one_electron(qf_engine* e, const double* charges, double* S, double* T, double* V)
{
    int natom, nshell, nprim, nbasis;
    const double *xyz, *exps, *coef;
    const int *s_atom, *s_lmin, *s_lmax, *s_nprim, *s_pstart, *s_fstart;
    int rc = qf_internal_basis(e, &natom, &xyz, &nshell, &s_atom, &s_lmin, &s_lmax, &s_nprim, &s_pstart, &s_fstart,
                               &nprim, &exps, &coef, &nbasis);
    if (rc) return rc;
    return qf_one_electron_ext(e, natom, xyz, charges, S, T, V);
}
