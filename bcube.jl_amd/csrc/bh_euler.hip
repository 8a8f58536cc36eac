// bh_euler.hip -- the residual of the reference's COVO Euler DG benchmark
// (benchmark/covo.jl:14-150): for a DG Q1 MultiFESpace (ρ, ρu, ρE, ϕ) on Quad4
// cells,
//   b = ∫ flux_Ω(u, v) dΩ - Σ_domains ∫ flux_roe ∘ (u⁻, u⁺, jump(v), n⁻) dΓ
// over interior faces and periodic boundary faces, and the explicit step
// u + Δt M⁻¹ b (covo.jl:31, forward_euler covo.jl:152).
//
// k_euler_faces computes the Roe flux at every face point once (one thread per
// point).  k_euler_dg then takes eight lanes per cell: lane g adds the face
// points q ≡ g (mod 2) of side g / 2 (that flux against the cell's own basis,
// with the sign of its role on the face) and the volume points q = g, g + 8,
// ... (w|J| ∇φ_l ⋅ F_k).  Both cells of a face read the same stored flux; each
// cell's 20 sums (5
// components x 4 basis functions) are reduced over the lanes in a fixed order
// and written by the cell that owns those dofs: no atomics, repeatable bit for
// bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bh_kernels.h"

namespace bh {
namespace {

constexpr int kEulerCellsPerBlock = 32;  // 8 lanes each, 256 threads
constexpr int kEulerMaxQ1 = 4;           // Gauss points per axis

// cell shapes: 0 = Quad4 (Q1, CGNS nodes), 1 = Tri3 (P1)
__device__ __forceinline__ int nsides_of(int ct) { return ct ? 3 : 4; }

// basis values at a reference point (4 slots; the 4th is 0 on triangles):
// Q1 tensor Lagrange on [-1,1]², first index fastest (hypercube.jl), or P1
// on the unit triangle λ = (1-ξ-η, ξ, η) (triangle.jl)
__device__ __forceinline__ void basis_at(int ct, double xi, double eta, double phi[4]) {
  if (ct) {
    phi[0] = 1.0 - xi - eta;
    phi[1] = xi;
    phi[2] = eta;
    phi[3] = 0.0;
  } else {
    const double a0 = 0.5 * (1.0 - xi), a1 = 0.5 * (1.0 + xi);
    const double b0 = 0.5 * (1.0 - eta), b1 = 0.5 * (1.0 + eta);
    phi[0] = a0 * b0;
    phi[1] = a1 * b0;
    phi[2] = a0 * b1;
    phi[3] = a1 * b1;
  }
}

// cell map Jacobian J[a][b] = ∂x_a/∂ξ_b at a reference point (Quad4 bilinear,
// Tri3 affine)
__device__ __forceinline__ void jacobian_at(int ct, const double X[4][2], double xi, double eta,
                                            double J[2][2]) {
  if (ct) {
    J[0][0] = X[1][0] - X[0][0];
    J[0][1] = X[2][0] - X[0][0];
    J[1][0] = X[1][1] - X[0][1];
    J[1][1] = X[2][1] - X[0][1];
    return;
  }
  const double dx[4] = {-(1.0 - eta) * 0.25, (1.0 - eta) * 0.25, (1.0 + eta) * 0.25,
                        -(1.0 + eta) * 0.25};
  const double de[4] = {-(1.0 - xi) * 0.25, -(1.0 + xi) * 0.25, (1.0 + xi) * 0.25,
                        (1.0 - xi) * 0.25};
  J[0][0] = J[0][1] = J[1][0] = J[1][1] = 0.0;
  for (int k = 0; k < 4; ++k) {
    J[0][0] += X[k][0] * dx[k];
    J[0][1] += X[k][0] * de[k];
    J[1][0] += X[k][1] * dx[k];
    J[1][1] += X[k][1] * de[k];
  }
}

// reference point of side s at Gauss abscissa t (mapping_face, mapping.jl:324-336):
// the side's corners (faces2nodes, shape.jl:253-254) from the first to the
// second, or backwards when flipped
__device__ __forceinline__ void side_point(int ct, int s, int flip, double t, double& xi,
                                           double& eta) {
  const double Vq[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
  const double Vt[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
  const int ns = nsides_of(ct);
  const int c0 = s, c1 = s + 1 == ns ? 0 : s + 1;
  const int a = flip ? c1 : c0, b = flip ? c0 : c1;
  const double* A = ct ? Vt[a] : Vq[a];
  const double* B = ct ? Vt[b] : Vq[b];
  xi = 0.5 * (1.0 - t) * A[0] + 0.5 * (1.0 + t) * B[0];
  eta = 0.5 * (1.0 - t) * A[1] + 0.5 * (1.0 + t) * B[1];
}

// reference outward normal of side s (shape.jl:277-278)
__device__ __forceinline__ void ref_normal(int ct, int s, double& nx, double& ny) {
  const double nq[4][2] = {{0.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}};
  const double r = 0.70710678118654752440;  // 1/√2
  const double nt[3][2] = {{0.0, -1.0}, {r, r}, {-1.0, 0.0}};
  nx = ct ? nt[s][0] : nq[s][0];
  ny = ct ? nt[s][1] : nq[s][1];
}

__device__ __forceinline__ double pressure(double rho, double rux, double ruy, double rE,
                                           double gamma) {
  // (γ-1)(ρE - tr(ρu ⊗ vel)/2), covo.jl:185-190
  const double vx = rux / rho, vy = ruy / rho;
  return (gamma - 1.0) * (rE - (rux * vx + ruy * vy) / 2.0);
}

// Roe's approximate Riemann flux across a face of unit normal n from the
// state uL (side⁻) to uR (side⁺), the flux of covo.jl:69-150 (no entropy
// fix).  Written as the textbook upwind-from-the-right form
//     F = F(uR)⋅n − Σ_k max(λ_k, 0) α_k K_k
// of Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics"
// (3rd ed., §11.2.1-11.2.2): Roe averages of (v_n, v_t, H), eigenvalues
// λ = v_n − a, v_n (entropy and shear waves), v_n + a, right eigenvectors K_k
// and wave strengths α_k of Δ = uR − uL, all in the face's frame (n, t) with
// t = (−n_y, n_x); the momentum rows are rotated back to (x, y) at the end.
// The sound speed is √|(γ−1)(H − |v|²/2)|, as the reference takes it.  The
// passive scalar ϕ is upwinded by the mean of the two normal velocities.
__device__ void roe_flux(const double uL[5], const double uR[5], double nx, double ny,
                         double gamma, double F[5]) {
  const double g1 = gamma - 1.0;
  // velocities in the face frame, pressure and total enthalpy of each state
  const double rL = uL[0], rR = uR[0];
  const double vxL = uL[1] / rL, vyL = uL[2] / rL, vxR = uR[1] / rR, vyR = uR[2] / rR;
  const double vnL = vxL * nx + vyL * ny, vtL = vyL * nx - vxL * ny;
  const double vnR = vxR * nx + vyR * ny, vtR = vyR * nx - vxR * ny;
  const double pL = pressure(rL, uL[1], uL[2], uL[3], gamma);
  const double pR = pressure(rR, uR[1], uR[2], uR[3], gamma);
  const double HL = (uL[3] + pL) / rL, HR = (uR[3] + pR) / rR;
  // Roe averages: weights √ρ
  const double sL = sqrt(rL), sR = sqrt(rR), inv = 1.0 / (sL + sR);
  const double vn = (sL * vnL + sR * vnR) * inv;
  const double vt = (sL * vtL + sR * vtR) * inv;
  const double H = (sL * HL + sR * HR) * inv;
  const double q2 = vn * vn + vt * vt;
  const double a = sqrt(fabs(g1 * (H - 0.5 * q2)));
  // jumps of the conserved variables (ρ, ρv_n, ρv_t, ρE)
  const double dr = rR - rL;
  const double dmn = rR * vnR - rL * vnL;
  const double dmt = rR * vtR - rL * vtL;
  const double dE = uR[3] - uL[3];
  // wave strengths (shear, entropy, the two acoustic waves)
  const double at = dmt - vt * dr;
  const double a2 = g1 / (a * a) * (dr * (H - vn * vn) + vn * dmn - (dE - at * vt));
  const double a1 = (dr * (vn + a) - dmn - a * a2) / (2.0 * a);
  const double a5 = dr - a1 - a2;
  const double w1 = fmax(vn - a, 0.0) * a1;  // λ⁺α of each wave
  const double w2 = fmax(vn, 0.0) * a2;
  const double wt = fmax(vn, 0.0) * at;
  const double w5 = fmax(vn + a, 0.0) * a5;
  // Σ λ⁺ α K in the face frame; K1 = (1, v_n − a, v_t, H − v_n a),
  // K2 = (1, v_n, v_t, |v|²/2), Kt = (0, 0, 1, v_t), K5 = (1, v_n + a, v_t, H + v_n a)
  const double s0 = w1 + w2 + w5;
  const double sn = w1 * (vn - a) + w2 * vn + w5 * (vn + a);
  const double st = (w1 + w2 + w5) * vt + wt;
  const double sE = w1 * (H - vn * a) + w2 * (0.5 * q2) + wt * vt + w5 * (H + vn * a);
  // the right state's physical flux, face frame
  const double mR = rR * vnR;
  const double fn = mR * vnR + pR - sn, ft = mR * vtR - st;
  F[0] = mR - s0;
  F[1] = fn * nx - ft * ny;
  F[2] = fn * ny + ft * nx;
  F[3] = mR * HR - sE;
  const double vbar = 0.5 * (vnL + vnR);
  F[4] = fmax(0.0, vbar) * uL[4] + fmin(0.0, vbar) * uR[4];
}

__device__ __forceinline__ void load_cell(const double* __restrict__ coords,
                                          const int32_t* __restrict__ c2n, int64_t c, int nn,
                                          double X[4][2]) {
  for (int k = 0; k < 4; ++k) {
    const int64_t n = c2n[4 * c + (k < nn ? k : 0)];
    X[k][0] = coords[2 * n];
    X[k][1] = coords[2 * n + 1];
  }
}

// the five components at one point: w[k] = Σ_{l < nl} φ_l u[maps[k][4c + l]]
__device__ __forceinline__ void point_state(const double* __restrict__ u,
                                            const int32_t* __restrict__ maps, int64_t stride,
                                            int64_t c, int nl, const double phi[4], double w[5]) {
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const int32_t* mk = maps + k * stride + 4 * c;
    double v = phi[0] * u[mk[0]] + phi[1] * u[mk[1]] + phi[2] * u[mk[2]];
    if (nl == 4) v += phi[3] * u[mk[3]];
    w[k] = v;
  }
}

// The Roe flux of every face point once: faces [nfaces][5] = (cell⁻, side⁻,
// cell⁺, side⁺, flip), G[(f*nq1 + q)*5 + k] = w_q |face|/2 F_k.
// SHAPES: 0 all Quad4, 1 all Tri3, 2 mixed (the cell types read per cell);
// uniform meshes fold every shape branch at compile time
template <int SHAPES>
__global__ __launch_bounds__(256) void k_euler_faces(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const int8_t* __restrict__ ctype, int64_t ncells, const int32_t* __restrict__ faces,
    int64_t nfaces, const int32_t* __restrict__ maps, EulerTabs T,
    const double* __restrict__ tabs, double gamma, const double* __restrict__ u,
    double* __restrict__ G) {
  const int64_t stride = 4 * ncells;
  const int nq1 = T.nq1;
  const double* gx = tabs + T.off_fx;
  const double* gw = tabs + T.off_fw;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nfaces * nq1;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = t / nq1;
    const int q = (int)(t - f * nq1);
    const int32_t* Fc = faces + 5 * f;
    const int64_t cm = Fc[0], cp = Fc[2];
    const int sm = Fc[1], sp = Fc[3], flip = Fc[4];
    const int ctm = SHAPES == 2 ? ctype[cm] : SHAPES, ctp = SHAPES == 2 ? ctype[cp] : SHAPES;
    double Xm[4][2];
    load_cell(coords, c2n, cm, ctm ? 3 : 4, Xm);
    const int sm1 = sm + 1 == nsides_of(ctm) ? 0 : sm + 1;
    const double ex = Xm[sm1][0] - Xm[sm][0], ey = Xm[sm1][1] - Xm[sm][1];
    const double dS = sqrt(ex * ex + ey * ey) / 2.0;
    double xm, em, xp, ep;
    side_point(ctm, sm, 0, gx[q], xm, em);
    side_point(ctp, sp, flip, gx[q], xp, ep);
    double J[2][2], phm[4], php[4];
    jacobian_at(ctm, Xm, xm, em, J);
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double r = 1.0 / det;
    // J⁻ᵀ n_ref, normalised (mapping.jl:843-855)
    double nrx, nry;
    ref_normal(ctm, sm, nrx, nry);
    const double a = nrx * (J[1][1] * r) + nry * (-J[1][0] * r);
    const double b = nrx * (-J[0][1] * r) + nry * (J[0][0] * r);
    const double inv = 1.0 / sqrt(a * a + b * b);
    basis_at(ctm, xm, em, phm);
    basis_at(ctp, xp, ep, php);
    double wm[5], wp[5], F[5];
    point_state(u, maps, stride, cm, ctm ? 3 : 4, phm, wm);
    point_state(u, maps, stride, cp, ctp ? 3 : 4, php, wp);
    roe_flux(wm, wp, a * inv, b * inv, gamma, F);
    const double wf = gw[q] * dS;
#pragma unroll
    for (int k = 0; k < 5; ++k) G[t * 5 + k] = wf * F[k];
  }
}

// side table: [ncells][4][4] = (face, 0, role, flip); role 0: no face term on
// this side, 1: this cell is the face's side⁻, 2: its side⁺.  Volume tables
// per shape (T.off_*[shape]): weights, basis values [np][4], reference
// gradients [np][4][2], geometric basis gradients [np][4][2].
template <int SHAPES>
__global__ __launch_bounds__(256) void k_euler_dg(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const int8_t* __restrict__ ctype, int64_t ncells, const int32_t* __restrict__ sides,
    const int32_t* __restrict__ maps, EulerTabs T, const double* __restrict__ tabs,
    double gamma, int terms, double fcoef, int step, double dt, const double* __restrict__ u,
    const double* __restrict__ G, double* __restrict__ out) {
  const int lc = threadIdx.x >> 3, g = threadIdx.x & 7;
  const int64_t c = (int64_t)blockIdx.x * kEulerCellsPerBlock + lc;
  const bool valid = c < ncells;
  const int64_t stride = 4 * ncells;
  const int nq1 = T.nq1;
  const double* gx = tabs + T.off_fx;
  const int ct = SHAPES == 2 ? (valid ? ctype[c] : 0) : SHAPES;
  const int nl = ct ? 3 : 4;
  double acc[5][4];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int l = 0; l < 4; ++l) acc[k][l] = 0.0;
  if (valid && (terms & 1)) {
    // volume points q = g, g + 8, ... of the shape's rule
    const double* w_ = tabs + T.off_w[ct];
    const double* ph_ = tabs + T.off_phi[ct];
    const double* dph_ = tabs + T.off_dphi[ct];
    const double* dg_ = tabs + T.off_dg[ct];
    const double* xq_ = tabs + T.off_xi[ct];
    for (int q = g; q < T.npts[ct]; q += 8) {
      double X[4][2];
      load_cell(coords, c2n, c, nl, X);
      double J[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
      double phi[4], dq[4][2];
      if constexpr (SHAPES == 0) {
        // all Quad4: the Q1 basis and the bilinear map at the point, in closed form
        const double xi = xq_[2 * q], eta = xq_[2 * q + 1];
        jacobian_at(0, X, xi, eta, J);
        basis_at(0, xi, eta, phi);
        const double a0 = 0.5 * (1.0 - xi), a1 = 0.5 * (1.0 + xi);
        const double b0 = 0.5 * (1.0 - eta), b1 = 0.5 * (1.0 + eta);
        dq[0][0] = -0.5 * b0; dq[0][1] = -0.5 * a0;
        dq[1][0] = 0.5 * b0;  dq[1][1] = -0.5 * a1;
        dq[2][0] = -0.5 * b1; dq[2][1] = 0.5 * a0;
        dq[3][0] = 0.5 * b1;  dq[3][1] = 0.5 * a1;
      } else {
        for (int k = 0; k < nl; ++k) {
          const double d0 = dg_[(q * 4 + k) * 2], d1 = dg_[(q * 4 + k) * 2 + 1];
          J[0][0] += X[k][0] * d0;
          J[0][1] += X[k][0] * d1;
          J[1][0] += X[k][1] * d0;
          J[1][1] += X[k][1] * d1;
        }
#pragma unroll
        for (int l = 0; l < 4; ++l) {
          phi[l] = ph_[q * 4 + l];
          dq[l][0] = dph_[(q * 4 + l) * 2];
          dq[l][1] = dph_[(q * 4 + l) * 2 + 1];
        }
      }
      const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double r = 1.0 / det;
      const double Ji[2][2] = {{J[1][1] * r, -J[0][1] * r}, {-J[1][0] * r, J[0][0] * r}};
      double w5[5];
      point_state(u, maps, stride, c, nl, phi, w5);
      // _flux_Ω (covo.jl:44-58)
      const double rho = w5[0], rux = w5[1], ruy = w5[2], rE = w5[3], pf = w5[4];
      const double vx = rux / rho, vy = ruy / rho;
      const double p = pressure(rho, rux, ruy, rE, gamma);
      const double F[5][2] = {{rux, ruy},
                              {rux * vx + p, rux * vy},
                              {ruy * vx, ruy * vy + p},
                              {(rE + p) * vx, (rE + p) * vy},
                              {pf * vx, pf * vy}};
      const double wd = w_[q] * fabs(det);
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        // ∇φ = J⁻ᵀ ∇̂φ
        const double d0 = dq[l][0], d1 = dq[l][1];
        const double gxl = d0 * Ji[0][0] + d1 * Ji[1][0];
        const double gyl = d0 * Ji[0][1] + d1 * Ji[1][1];
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k][l] += wd * (gxl * F[k][0] + gyl * F[k][1]);
      }
    }
  }
  if (valid && (terms & 2)) {
    // side s = g / 2, its face points q = g % 2, g % 2 + 2, ...: the face's
    // flux times this cell's basis at the point, with the sign of its role
    const int s = g >> 1;
    const int32_t* S = sides + (4 * c + s) * 4;
    const int role = S[2];
    if (role != 0) {
      const int64_t f = S[0];
      const int flip = role == 1 ? 0 : S[3];
      // fcoef ∫ F ⋅ jump(v), jump(v) = v⁻ - v⁺ (fcoef = -1 in covo.jl's l)
      const double sg = role == 1 ? fcoef : -fcoef;
      for (int q = g & 1; q < nq1; q += 2) {
        double xi, eta, ph[4];
        side_point(ct, s, flip, gx[q], xi, eta);
        basis_at(ct, xi, eta, ph);
        const double* Gq = G + (f * nq1 + q) * 5;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const double Fk = Gq[k];
#pragma unroll
          for (int l = 0; l < 4; ++l) acc[k][l] += sg * (Fk * ph[l]);
        }
      }
    }
  }
  // sums over the cell's 8 lanes, butterfly in a fixed order (a + b = b + a:
  // all 8 lanes hold the same totals)
#pragma unroll
  for (int off = 1; off < 8; off <<= 1)
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int l = 0; l < 4; ++l) acc[k][l] += __shfl_xor(acc[k][l], off);
  if (!valid) return;
  if (!step) {
    // b += l(v) (assemble_linear!): lane g writes entries g, g + 8, g + 16
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int l = 0; l < 4; ++l)
        if (((k * 4 + l) & 7) == g && l < nl) out[maps[k * stride + 4 * c + l]] += acc[k][l];
    return;
  }
  if (g >= 5) return;
  // u + Δt M⁻¹ b, one component per lane
  double X[4][2];
  load_cell(coords, c2n, c, nl, X);
  double rhs[4];
#pragma unroll
  for (int k = 0; k < 5; ++k)
    if (k == g)
#pragma unroll
      for (int l = 0; l < 4; ++l) rhs[l] = acc[k][l];
  const double* w_ = tabs + T.off_w[ct];
  const double* ph_ = tabs + T.off_phi[ct];
  const double* dg_ = tabs + T.off_dg[ct];
  const double* xq_ = tabs + T.off_xi[ct];
  auto wdet = [&](int q) {
    double J[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if constexpr (SHAPES == 0) {
      jacobian_at(0, X, xq_[2 * q], xq_[2 * q + 1], J);
      return w_[q] * fabs(J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    }
    for (int k = 0; k < nl; ++k) {
      const double d0 = dg_[(q * 4 + k) * 2], d1 = dg_[(q * 4 + k) * 2 + 1];
      J[0][0] += X[k][0] * d0;
      J[0][1] += X[k][0] * d1;
      J[1][0] += X[k][1] * d0;
      J[1][1] += X[k][1] * d1;
    }
    return w_[q] * fabs(J[0][0] * J[1][1] - J[0][1] * J[1][0]);
  };
  if (ct == 0 && T.off_pinv >= 0) {
    // as many points as basis functions (2x2 Gauss): M = Φᵀ D Φ with Φ[q][l]
    // = φ_l(ξ_q) square and D = diag(w_q |J(ξ_q)|), so M⁻¹ = Φ⁻¹ D⁻¹ Φ⁻ᵀ
    const double* Pi = tabs + T.off_pinv;
    double y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      y[q] = (Pi[q] * rhs[0] + Pi[4 + q] * rhs[1] + Pi[8 + q] * rhs[2] + Pi[12 + q] * rhs[3]) /
             wdet(q);
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const double x = Pi[4 * l] * y[0] + Pi[4 * l + 1] * y[1] + Pi[4 * l + 2] * y[2] +
                       Pi[4 * l + 3] * y[3];
      const int32_t id = maps[g * stride + 4 * c + l];
      out[id] = u[id] + dt * x;
    }
    return;
  }
  // the cell's mass matrix (same rule), Gaussian elimination (SPD)
  double M[4][4] = {};
  for (int q = 0; q < T.npts[ct]; ++q) {
    const double wd = wdet(q);
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nl; ++j) M[i][j] += wd * ph_[q * 4 + i] * ph_[q * 4 + j];
  }
  for (int i = 0; i < nl; ++i) {
    const double piv = 1.0 / M[i][i];
    for (int r = i + 1; r < nl; ++r) {
      const double fct = M[r][i] * piv;
      for (int j = i; j < nl; ++j) M[r][j] -= fct * M[i][j];
      rhs[r] -= fct * rhs[i];
    }
  }
  for (int i = nl - 1; i >= 0; --i) {
    double v = rhs[i];
    for (int j = i + 1; j < nl; ++j) v -= M[i][j] * rhs[j];
    rhs[i] = v / M[i][i];
  }
  for (int l = 0; l < nl; ++l) {
    const int32_t id = maps[g * stride + 4 * c + l];
    out[id] = u[id] + dt * rhs[l];
  }
}

}  // namespace

hipError_t launch_euler_dg(hipStream_t s, const double* coords, const int32_t* c2n,
                           const int8_t* ctype, int shapes, int64_t ncells, const int32_t* sides,
                           const int32_t* faces, int64_t nfaces, const int32_t* maps,
                           const EulerTabs& T, const double* tabs, double gamma, int terms,
                           double fcoef, int step, double dt, const double* u, double* G,
                           double* out) {
  if (ncells <= 0) return hipSuccess;
  if (shapes < 0 || shapes > 2) return hipErrorInvalidValue;
  if (T.nq1 < 1 || T.nq1 > kEulerMaxQ1) return hipErrorInvalidValue;
  const int64_t nb = (ncells + kEulerCellsPerBlock - 1) / kEulerCellsPerBlock;
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  if ((terms & 2) && nfaces > 0) {
    const int64_t n = nfaces * T.nq1;
    const dim3 grid((unsigned)((n + 255) / 256 > 65536 ? 65536 : (n + 255) / 256));
#define BH_EF(SH)                                                                               \
  hipLaunchKernelGGL(k_euler_faces<SH>, grid, dim3(256), 0, s, coords, c2n, ctype, ncells, faces, \
                     nfaces, maps, T, tabs, gamma, u, G)
    if (shapes == 0) BH_EF(0);
    else if (shapes == 1) BH_EF(1);
    else BH_EF(2);
#undef BH_EF
  }
#define BH_EC(SH)                                                                               \
  hipLaunchKernelGGL(k_euler_dg<SH>, dim3((unsigned)nb), dim3(8 * kEulerCellsPerBlock), 0, s,      \
                     coords, c2n, ctype, ncells, sides, maps, T, tabs, gamma, terms, fcoef, step, \
                     dt, u, G, out)
  if (shapes == 0) BH_EC(0);
  else if (shapes == 1) BH_EC(1);
  else BH_EC(2);
#undef BH_EC
  return hipGetLastError();
}

}  // namespace bh
