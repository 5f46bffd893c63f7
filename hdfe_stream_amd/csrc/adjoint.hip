// Row passes of the adjoint (backward) of a fitted fixed-effect regression,
// for gfx950.
//
// The fit's coefficients theta (covariates and every level's effect) solve
// Z'WZ theta = Z'W y.  For a scalar loss with gradient g = dL/dtheta, one
// adjoint solve u = (Z'WZ)^- g gives every row's sensitivity z_i'u, and
//   dL/dy_i = w_i z_i'u,     dL/dw_i = e_i z_i'u (+ the loss's own w_i term),
// e the residuals.  With covariates X beside the fixed effects F, the Schur
// block on X gives u_X = A_xx^-1 (g_X - X'W F T^- g_F) and z_i'u =
// (F T^- g_F)_i + x~_i'u_X, x~ = M_F x the partialled covariates and A_xx =
// X'W M_F X (the fit's V'M_D V block).  The reference has no autograd; its
// adjoint hook is solve_normal_equations / predict (normal_equations.py:
// 65-231), whose z'u_j and w_i z_i'u_j are these row values.
//
// Reductions are fixed row blocks added in a fixed order.
#include <vector>

#include "devbuf.h"

namespace hdfe {

struct CompCoef {
    double a_psi, a_alpha, a_cov;   // the loss's weights on var(psi), var(alpha), cov
    double psibar, alphabar;        // weighted means of the effects over rows
    double Dp, W;                   // the divisor W (R - 1) / R and the weight total
    double Vpsi, Valpha, Cov;       // the three components
};

// Per row: r_psi = dL/dpsi_i / w_i and r_alpha = dL/dalpha_i / w_i of the
// plug-in components (leaveout.py:807-832), and the components' own
// derivative in w_i at fixed effects.
__global__ void k_adj_component_rows(int64_t G, const int *__restrict__ gstart,
                                     const int *__restrict__ psi_codes, int64_t psi_off,
                                     const double *__restrict__ gam,
                                     const double *__restrict__ geff, CompCoef c,
                                     double *__restrict__ r_psi, double *__restrict__ r_alpha,
                                     double *__restrict__ direct) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < G;
         g += (int64_t)gridDim.x * blockDim.x) {
        const double dal = geff[g] - c.alphabar;
        for (int i = gstart[g]; i < gstart[g + 1]; ++i) {
            const double dps = gam[psi_off + psi_codes[i]] - c.psibar;
            r_psi[i] = (2.0 * c.a_psi * dps + c.a_cov * dal) / c.Dp;
            r_alpha[i] = (2.0 * c.a_alpha * dal + c.a_cov * dps) / c.Dp;
            direct[i] = c.a_psi * (dps * dps / c.Dp - c.Vpsi / c.W) +
                        c.a_alpha * (dal * dal / c.Dp - c.Valpha / c.W) +
                        c.a_cov * (dps * dal / c.Dp - c.Cov / c.W);
        }
    }
}

// part[b, j] = sum over block b's rows of w_i V[i, xcols[j]] R_i.
__global__ void k_adj_xwr(int64_t n, int m, const double *__restrict__ V,
                          const int *__restrict__ xcols, int kx, const double *__restrict__ w,
                          const double *__restrict__ R, double *__restrict__ part) {
    __shared__ double lds[4];
    const int64_t r0 = part_begin(n, blockIdx.x, gridDim.x), r1 = part_begin(n, blockIdx.x + 1, gridDim.x);
    for (int j = 0; j < kx; ++j) {
        const int col = xcols[j];
        double s = 0.0;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
            s += (w ? w[i] : 1.0) * V[i * m + col] * R[i];
        s = block_sum<double>(s, lds);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * kx + j] = s;
    }
}

struct UX {
    double v[16];
};

// zu_i = rows_i + sum_j Xt[i, j] uX_j; gy = w zu; gw = e zu + direct.
__global__ void k_adj_gradients(int64_t n, int kx, const double *__restrict__ rows,
                                const double *__restrict__ Xt, UX u,
                                const double *__restrict__ w, const double *__restrict__ e,
                                const double *__restrict__ direct, double *__restrict__ gy,
                                double *__restrict__ gw) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double zu = rows ? rows[i] : 0.0;
        for (int j = 0; j < kx; ++j) zu += Xt[i * kx + j] * u.v[j];
        gy[i] = (w ? w[i] : 1.0) * zu;
        gw[i] = e[i] * zu + (direct ? direct[i] : 0.0);
    }
}

}  // namespace hdfe

using namespace hdfe;

extern "C" {

// r_psi, r_alpha, direct (n, device): see k_adj_component_rows; coef (host,
// 10) = a_psi, a_alpha, a_cov, psibar, alphabar, Dp, W, Vpsi, Valpha, Cov.
int hdfe_adj_component_rows(int64_t G, const int *gstart, const int *psi_codes, int64_t psi_off,
                            const double *gam, const double *geff, const double *coef,
                            double *r_psi, double *r_alpha, double *direct) {
    ARG_CHECK(gstart && psi_codes && gam && geff && coef && r_psi && r_alpha && direct,
              "hdfe_adj_component_rows: null argument");
    CompCoef c{coef[0], coef[1], coef[2], coef[3], coef[4], coef[5], coef[6], coef[7], coef[8],
               coef[9]};
    ARG_CHECK(c.Dp > 0.0 && c.W > 0.0, "hdfe_adj_component_rows: no weight");
    hipLaunchKernelGGL(k_adj_component_rows, dim3(grid1(G)), dim3(256), 0, 0, G, gstart, psi_codes,
                       psi_off, gam, geff, c, r_psi, r_alpha, direct);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

// out (host, kx) = X'W R for the covariate columns xcols (host) of V (n, m).
int hdfe_adj_xwr(int64_t n, int m, const double *V, const int *xcols, int kx, const double *w,
                 const double *R, double *out) {
    ARG_CHECK(V && xcols && R && out && kx >= 1, "hdfe_adj_xwr: bad argument");
    const int nb = 256;
    DevBuf<int> xc;
    DEV_ALLOC(xc, kx, nullptr);
    HIP_TRY(hipMemcpy(xc.p, xcols, kx * 4, hipMemcpyHostToDevice));
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * kx, nullptr);
    hipLaunchKernelGGL(k_adj_xwr, dim3(nb), dim3(256), 0, 0, n, m, V, xc.p, kx, w, R, part.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> h;
    HIP_TRY(to_host(h, part.p, (int64_t)nb * kx, 0));
    for (int j = 0; j < kx; ++j) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += h[(size_t)b * kx + j];
        out[j] = s;
    }
    return HDFE_STATUS_OK;
}

// gy, gw (n, device) from the rows of F T^- g_F (NULL: zero), the
// partialled covariates Xt (n, kx), u_X (host, kx <= 16), the residuals and
// the loss's own w term (NULL: none).
int hdfe_adj_gradients(int64_t n, int kx, const double *rows, const double *Xt, const double *uX,
                       const double *w, const double *e, const double *direct, double *gy,
                       double *gw) {
    ARG_CHECK(e && gy && gw && kx >= 0 && kx <= 16 && (kx == 0 || (Xt && uX)),
              "hdfe_adj_gradients: bad argument");
    UX u;
    for (int j = 0; j < 16; ++j) u.v[j] = j < kx ? uX[j] : 0.0;
    hipLaunchKernelGGL(k_adj_gradients, dim3(grid1(n)), dim3(256), 0, 0, n, kx, rows, Xt, u, w, e,
                       direct, gy, gw);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return HDFE_STATUS_OK;
}

}  // extern "C"
