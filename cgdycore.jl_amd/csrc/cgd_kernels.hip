// HIP kernels (gfx950) for the spectral-element tendency, the column
// Jacobian / Schur solve and the Rosenbrock-W stage arithmetic.
//
// Data layout in HBM (all level-fastest, so a wavefront's lanes read
// consecutive levels of one column):
//   state / caches   X[v][g][iz]                  (reference (Nz,NumG,NumV))
//   metric (pass 2)  met[e][c][ID][iz], c < 17    (built once from dXdxI/J/X)
//   metric (pass 1)  met1[e][c][ID][iz], c < 5    (S11,S12,S21,S22,J1+J2)
//   DSS mass         Ms[g][iz] = M1 + M2, M[K][g][iz]
//
// Element kernels: one thread per (node, level) point of one element and
// one chunk of <= 40 levels; thread t -> node = t / nzt, level = t % nzt.
// Horizontal derivatives exchange values through LDS tiles
// tile[f][node][level]; vertical neighbours are read directly (the same
// cache lines the neighbouring lanes load).  The DSS is a non-atomic
// read-modify-write: one launch per element colour (no two elements of a
// colour share a global node), so every global node accumulates its element
// contributions in a fixed order.
#include <hip/hip_runtime.h>

#include <stdint.h>

#define CGD_NMAX 5
#define CGD_NMET 17  // pass-2 metric components
#define CGD_NMET1 5  // pass-1 metric components

struct KArgs {
  int Nz, NumG, nzt;
  double Grav, Omega, Rd, Cpd, p0, kappa;
  double Kc, Kg, Kd;
  double D[CGD_NMAX * CGD_NMAX];   // D[i*N + k] = DS[i,k]
  double DWm[CGD_NMAX * CGD_NMAX]; // DW[i*N + k]
  const int *glob;                 // [NF][NN], 0-based
  const double *met;               // [NF][17][NN][Nz]
  const double *met1;              // [NF][5][NN][Nz]
  const double *Ms;                // [NumG][Nz]
  const double *M;                 // [2][NumG][Nz]
};

// metric component indices of met (dXdxI[a,b,K] -> dabK, 1-based a,b,K)
enum {
  M111 = 0, M112, M121, M122, M211, M212, M221, M222,
  M311, M312, M321, M322, M331, M332, MJ1, MJ2, MCOR
};

// ------------------------------------------------------------------------
// Equation of state (src/Models/Equation.jl:25-38,79)
__device__ __forceinline__ double eos_p(const KArgs &a, double rhoth) {
  const double y = 1.0 / (1.0 - a.kappa);
  return a.p0 * exp(y * log(a.Rd * rhoth / a.p0));
}
__device__ __forceinline__ double eos_dpdrhoth(double Rd, double p0, double kappa, double rhoth) {
  return Rd * pow(Rd * rhoth / p0, kappa / (1.0 - kappa));
}

// Julia min/max on non-NaN values
__device__ __forceinline__ double jmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double jmax(double a, double b) { return (b > a) ? b : a; }

// RecU4 (src/GPU/OperatorKernel.jl:1044-1059)
__device__ __forceinline__ void recu4(double cLL, double cL, double cR, double cRR, double JLL, double JL,
                                      double JR, double JRR, double &cFL, double &cFR) {
  double kR = (JL / (JL + JR)) * ((JLL + JL) / (JLL + JL + JR));
  double kL = (JL / (JLL + JL)) * (JR / (JLL + JL + JR));
  double r = (cL - cLL) / (cR - cL + 1.e-20);
  cFL = cL + jmax(0.0, jmin(r, jmin(kL * r + kR, 1.0))) * (cR - cL);
  kL = (JR / (JR + JL)) * ((JRR + JR) / (JL + JR + JRR));
  kR = (JR / (JRR + JR)) * (JL / (JL + JR + JRR));
  r = (cR - cRR) / (cL - cR + 1.e-20);
  cFR = cR + jmax(0.0, jmin(r, jmin(kR * r + kL, 1.0))) * (cL - cR);
}

// ------------------------------------------------------------------------
// LDS tile helpers
template <int N> struct Tiles {
  double *t;  // tile base
  int nzt;
  __device__ __forceinline__ double &at(int f, int node, int l) const { return t[(f * N * N + node) * nzt + l]; }
};

// sum_k D[I,k] f(k,J)
template <int N>
__device__ __forceinline__ double dx(const double *Dm, const Tiles<N> &T, int f, int I, int J, int l) {
  double s = Dm[I * N] * T.at(f, J * N, l);
#pragma unroll
  for (int k = 1; k < N; ++k) s += Dm[I * N + k] * T.at(f, k + J * N, l);
  return s;
}
// sum_k D[J,k] f(I,k)
template <int N>
__device__ __forceinline__ double dy(const double *Dm, const Tiles<N> &T, int f, int I, int J, int l) {
  double s = Dm[J * N] * T.at(f, I, l);
#pragma unroll
  for (int k = 1; k < N; ++k) s += Dm[J * N + k] * T.at(f, I + k * N, l);
  return s;
}
// sum_k (D[I,k] fx(k,J) + D[J,k] fy(I,k)), pairwise as in the reference loops
template <int N>
__device__ __forceinline__ double dxy(const double *Dm, const Tiles<N> &T, int fx, int fy, int I, int J, int l) {
  double s = Dm[I * N] * T.at(fx, J * N, l) + Dm[J * N] * T.at(fy, I, l);
#pragma unroll
  for (int k = 1; k < N; ++k) s += Dm[I * N + k] * T.at(fx, k + J * N, l) + Dm[J * N + k] * T.at(fy, I + k * N, l);
  return s;
}
// acc + sum_k (D[I,k] fx(k,J) + D[J,k] fy(I,k)), accumulated term by term
template <int N>
__device__ __forceinline__ double dxy_acc(double s, const double *Dm, const Tiles<N> &T, int fx, int fy, int I,
                                          int J, int l) {
#pragma unroll
  for (int k = 0; k < N; ++k) s += Dm[I * N + k] * T.at(fx, k + J * N, l) + Dm[J * N + k] * T.at(fy, I + k * N, l);
  return s;
}
// sum_k (D[I,k] fx(k,J) cx(k,J) + D[J,k] fy(I,k) cy(I,k))
template <int N>
__device__ __forceinline__ double dxy_prod(const double *Dm, const Tiles<N> &T, int fx, int fy, int fc, int I, int J,
                                           int l) {
  double s = Dm[I * N] * T.at(fx, J * N, l) * T.at(fc, J * N, l) + Dm[J * N] * T.at(fy, I, l) * T.at(fc, I, l);
#pragma unroll
  for (int k = 1; k < N; ++k)
    s += Dm[I * N + k] * T.at(fx, k + J * N, l) * T.at(fc, k + J * N, l) +
         Dm[J * N + k] * T.at(fy, I + k * N, l) * T.at(fc, I + k * N, l);
  return s;
}

template <int N> __device__ __forceinline__ void load_dmats(const KArgs &a, double *Ds, double *DWs) {
  const int t = threadIdx.x;
  if (t < N * N) {
    Ds[t] = a.D[t];
    DWs[t] = a.DWm[t];
  }
}

// ------------------------------------------------------------------------
// Pass 1: HyperViscKernel! (src/GPU/ViscKernel.jl:1-84).
// Grid: one workgroup per (element of the colour, level chunk).
template <int N>
__global__ void __launch_bounds__(1024) k_fcn1(KArgs a, const int *__restrict__ elist, int nchunk,
                                               const double *__restrict__ U, double *__restrict__ T1) {
  constexpr int NN = N * N;
  extern __shared__ double smem[];
  double *Ds = smem, *DWs = smem + NN;
  Tiles<N> T{smem + 2 * NN, a.nzt};
  load_dmats<N>(a, Ds, DWs);
  const int Nz = a.Nz, nzt = a.nzt;
  const int e = elist[blockIdx.x / nchunk];
  const int chunk = blockIdx.x - (blockIdx.x / nchunk) * nchunk;
  const int t = threadIdx.x;
  const int node = t / nzt, l = t - node * nzt;
  const int iz = chunk * nzt + l;
  const bool in_tile = node < NN;
  const bool valid = in_tile && iz < Nz;
  const int I = node % N, J = node / N;
  const size_t vs = (size_t)a.NumG * Nz;
  const size_t NNZ = (size_t)NN * Nz;

  double S11 = 0, S12 = 0, S21 = 0, S22 = 0, Js = 1, Ms = 1;
  size_t ui = 0;
  double uC = 0, vC = 0, uD = 0, vD = 0, Th = 0;
  if (valid) {
    const int g = a.glob[e * NN + node];
    ui = (size_t)g * Nz + iz;
    const size_t mb = ((size_t)e * CGD_NMET1 * NN + node) * Nz + iz;
    S11 = a.met1[mb];
    S12 = a.met1[mb + NNZ];
    S21 = a.met1[mb + 2 * NNZ];
    S22 = a.met1[mb + 3 * NNZ];
    Js = a.met1[mb + 4 * NNZ];
    Ms = a.Ms[ui];
    const double rho = U[ui], u = U[ui + vs], v = U[ui + 2 * vs], rth = U[ui + 4 * vs];
    uC = S11 * v - S12 * u;  // Curl12
    vC = S21 * v - S22 * u;
    uD = S11 * u + S12 * v;  // Contra12
    vD = S21 * u + S22 * v;
    Th = rth / rho;
  }
  if (in_tile) {
    T.at(0, node, l) = uC;
    T.at(1, node, l) = vC;
    T.at(2, node, l) = uD;
    T.at(3, node, l) = vD;
    T.at(4, node, l) = Th;
  }
  __syncthreads();
  double Curl = 0, Div = 0, ThCx = 0, ThCy = 0;
  if (valid) {
    const double Dxc = dx<N>(Ds, T, 4, I, J, l);
    const double Dyc = dy<N>(Ds, T, 4, I, J, l);
    Curl = dxy<N>(Ds, T, 0, 1, I, J, l) / Js;
    Div = dxy<N>(Ds, T, 2, 3, I, J, l) / Js;
    const double GDx = (S11 * Dxc + S21 * Dyc) / Js;  // Grad12 with J
    const double GDy = (S12 * Dxc + S22 * Dyc) / Js;
    ThCx = S11 * GDx + S12 * GDy;  // Contra12
    ThCy = S21 * GDx + S22 * GDy;
  }
  __syncthreads();
  if (in_tile) {
    T.at(0, node, l) = Curl;
    T.at(1, node, l) = Div;
    T.at(2, node, l) = ThCx;
    T.at(3, node, l) = ThCy;
  }
  __syncthreads();
  if (valid) {
    const double DxCurl = dx<N>(DWs, T, 0, I, J, l);
    const double DyCurl = dy<N>(DWs, T, 0, I, J, l);
    const double DxDiv = dx<N>(DWs, T, 1, I, J, l);
    const double DyDiv = dy<N>(DWs, T, 1, I, J, l);
    const double DivTh = dxy<N>(DWs, T, 2, 3, I, J, l);
    const double FuC = S12 * DxCurl + S22 * DyCurl;  // Rot12
    const double FvC = -S11 * DxCurl - S21 * DyCurl;
    const double FuD = S11 * DxDiv + S21 * DyDiv;  // Grad12
    const double FvD = S12 * DxDiv + S22 * DyDiv;
    T1[ui] += -FuC / Ms;
    T1[ui + vs] += -FvC / Ms;
    T1[ui + 2 * vs] += FuD / Ms;
    T1[ui + 3 * vs] += FvD / Ms;
    T1[ui + 4 * vs] += DivTh / Ms;
  }
}

// ------------------------------------------------------------------------
// Pass 2: HyperViscKoeffKernel! + MomentumVectorInvariantCoriolisKernel! +
// GradFullKernel! + DivRhoTrUpwind3New2Kernel! fused into one element pass
// (src/GPU/ViscKernel.jl:86-167, src/GPU/MomentumKernel.jl:1-178,
// src/GPU/GradKineticKernel.jl:1-129, src/GPU/OperatorKernel.jl:637-720).
// Launched once per (colour, chunk) so that the F4 contribution of a
// chunk's first level to the interface below it is ordered after the
// previous chunk of the same element.
template <int N>
__global__ void __launch_bounds__(1024) k_fcn2(KArgs a, const int *__restrict__ elist, int chunk,
                                               const double *__restrict__ U, const double *__restrict__ C,
                                               const double *__restrict__ P, double *__restrict__ F) {
  constexpr int NN = N * N;
  extern __shared__ double smem[];
  double *Ds = smem, *DWs = smem + NN;
  Tiles<N> T{smem + 2 * NN, a.nzt};
  load_dmats<N>(a, Ds, DWs);
  const int Nz = a.Nz, nzt = a.nzt;
  const int e = elist[blockIdx.x];
  const int t = threadIdx.x;
  const int node = t / nzt, l = t - node * nzt;
  const int iz = chunk * nzt + l;
  const bool in_tile = node < NN;
  const bool valid = in_tile && iz < Nz;
  const int I = node % N, J = node / N;
  const size_t vs = (size_t)a.NumG * Nz;
  const size_t NNZ = (size_t)NN * Nz;
  const bool has_lo = iz > 0, has_hi = iz < Nz - 1;

  // ---- loads ----
  double rho = 1, u = 0, v = 0, w = 0, rth = 1, p = 0;
  double c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
  double d[CGD_NMET];
#pragma unroll
  for (int c = 0; c < CGD_NMET; ++c) d[c] = 1.0;
  double Ms = 1, M1 = 1, M2 = 1;
  // level iz-1
  double rho_m = 1, u_m = 0, v_m = 0, w_m = 0, rth_m = 1, M2_m = 1, d312_m = 0, d322_m = 0, d332_m = 1, Jc_m = 1;
  // level iz-2 (clamped)
  double rho_mm = 1, rth_mm = 1, Jc_mm = 1;
  // level iz+1
  double rho_p = 1, u_p = 0, v_p = 0, rth_p = 1, p_p = 0, M1_p = 1, J1_p = 1, Jc_p = 1;
  double d311_p = 0, d321_p = 0, d331_p = 1;
  // level iz+2 (clamped)
  double rho_pp = 1, rth_pp = 1, Jc_pp = 1;
  size_t ui = 0;
  if (valid) {
    const int g = a.glob[e * NN + node];
    ui = (size_t)g * Nz + iz;
    const size_t mb = ((size_t)e * CGD_NMET * NN + node) * Nz + iz;
    rho = U[ui];
    u = U[ui + vs];
    v = U[ui + 2 * vs];
    w = U[ui + 3 * vs];
    rth = U[ui + 4 * vs];
    c1 = C[ui];
    c2 = C[ui + vs];
    c3 = C[ui + 2 * vs];
    c4 = C[ui + 3 * vs];
    c5 = C[ui + 4 * vs];
    p = P[ui];
#pragma unroll
    for (int c = 0; c < CGD_NMET; ++c) d[c] = a.met[mb + c * NNZ];
    Ms = a.Ms[ui];
    M1 = a.M[ui];
    M2 = a.M[ui + vs];
    if (has_lo) {
      const size_t um = ui - 1, mm = mb - 1;
      rho_m = U[um];
      u_m = U[um + vs];
      v_m = U[um + 2 * vs];
      w_m = U[um + 3 * vs];
      rth_m = U[um + 4 * vs];
      M2_m = a.M[um + vs];
      d312_m = a.met[mm + M312 * NNZ];
      d322_m = a.met[mm + M322 * NNZ];
      d332_m = a.met[mm + M332 * NNZ];
      Jc_m = a.met[mm + MJ1 * NNZ] + a.met[mm + MJ2 * NNZ];
      const int dm2 = iz >= 2 ? 2 : 1;
      rho_mm = U[ui - dm2];
      rth_mm = U[ui - dm2 + 4 * vs];
      Jc_mm = a.met[mb - dm2 + MJ1 * NNZ] + a.met[mb - dm2 + MJ2 * NNZ];
    }
    if (has_hi) {
      const size_t up = ui + 1, mp = mb + 1;
      rho_p = U[up];
      u_p = U[up + vs];
      v_p = U[up + 2 * vs];
      rth_p = U[up + 4 * vs];
      p_p = P[up];
      M1_p = a.M[up];
      J1_p = a.met[mp + MJ1 * NNZ];
      Jc_p = J1_p + a.met[mp + MJ2 * NNZ];
      d311_p = a.met[mp + M311 * NNZ];
      d321_p = a.met[mp + M321 * NNZ];
      d331_p = a.met[mp + M331 * NNZ];
      const int dp2 = iz + 2 <= Nz - 1 ? 2 : 1;
      rho_pp = U[ui + dp2];
      rth_pp = U[ui + dp2 + 4 * vs];
      Jc_pp = a.met[mb + dp2 + MJ1 * NNZ] + a.met[mb + dp2 + MJ2 * NNZ];
    }
  }
  const double S11 = d[M111] + d[M112], S12 = d[M121] + d[M122];
  const double S21 = d[M211] + d[M212], S22 = d[M221] + d[M222];
  const double Js = d[MJ1] + d[MJ2];
  // free-slip bottom (MomentumKernel.jl:35-36)
  const double w_lo = has_lo ? w_m : -(d[M311] * u + d[M321] * v) / d[M331];
  const double w_hi = w;

  // ---- phase 1: hyperviscosity inputs, kinetic energy, pressure ----
  const double Kh = 0.5 * (u * u + v * v);
  const double K1 = Kh + 0.5 * (w_lo * w_lo);
  const double K2 = Kh + 0.5 * (w_hi * w_hi);
  if (in_tile) {
    T.at(0, node, l) = S11 * c2 - S12 * c1;  // Curl12(Cache1, Cache2)
    T.at(1, node, l) = S21 * c2 - S22 * c1;
    T.at(2, node, l) = S11 * c3 + S12 * c4;  // Contra12(Cache3, Cache4)
    T.at(3, node, l) = S21 * c3 + S22 * c4;
    T.at(4, node, l) = c5;
    T.at(5, node, l) = K1;
    T.at(6, node, l) = K2;
    T.at(7, node, l) = p;
  }
  __syncthreads();
  double Curl = 0, Div = 0, ThCx = 0, ThCy = 0;
  double f2 = 0, f3 = 0, f4up = 0, f4lo = 0;
  const double den_up = rho_p * M1_p + rho * M2;   // F4[iz] mass
  const double den_lo = rho_m * M2_m + rho * M1;   // F4[iz-1] mass
  if (valid) {
    const double Dxc = dx<N>(Ds, T, 4, I, J, l);
    const double Dyc = dy<N>(Ds, T, 4, I, J, l);
    Curl = dxy<N>(Ds, T, 0, 1, I, J, l) / Js;
    Div = dxy<N>(Ds, T, 2, 3, I, J, l) / Js;
    const double GDx = (S11 * Dxc + S21 * Dyc) / Js;
    const double GDy = (S12 * Dxc + S22 * Dyc) / Js;
    ThCx = rho * (S11 * GDx + S12 * GDy);  // Contra12(U1, ...)
    ThCy = rho * (S21 * GDx + S22 * GDy);
    // GradFullKernel!
    const double DXK1 = dx<N>(Ds, T, 5, I, J, l), DYK1 = dy<N>(Ds, T, 5, I, J, l);
    const double DXK2 = dx<N>(Ds, T, 6, I, J, l), DYK2 = dy<N>(Ds, T, 6, I, J, l);
    const double DXp = dx<N>(Ds, T, 7, I, J, l), DYp = dy<N>(Ds, T, 7, I, J, l);
    double GuF1 = -(d[M111] * DXK1 + d[M211] * DYK1);
    double GvF1 = -(d[M121] * DXK1 + d[M221] * DYK1);
    double GuF2 = -(d[M112] * DXK2 + d[M212] * DYK2);
    double GvF2 = -(d[M122] * DXK2 + d[M222] * DYK2);
    double Gu = S11 * DXp + S21 * DYp;
    double Gv = S12 * DXp + S22 * DYp;
    double GwF1 = 0, GwF2 = 0, GradZ;
    if (has_lo) {
      const double K2_m = 0.5 * (u_m * u_m + v_m * v_m) + 0.5 * (w_m * w_m);
      GradZ = 0.5 * (K1 - K2_m);
      GuF1 += -GradZ * d[M311];
      GvF1 += -GradZ * d[M321];
      GwF1 += -rho * GradZ * d[M331];
    }
    if (has_hi) {
      const double K1_p = 0.5 * (u_p * u_p + v_p * v_p) + 0.5 * (w * w);
      GradZ = 0.5 * (K1_p - K2);
      GuF2 += -GradZ * d[M312];
      GvF2 += -GradZ * d[M322];
      GwF2 += -rho * GradZ * d[M332];
    }
    GradZ = 0.5 * (K2 - K1);
    GuF2 += -GradZ * d[M312];
    GuF1 += -GradZ * d[M311];
    GvF2 += -GradZ * d[M322];
    GvF1 += -GradZ * d[M321];
    GwF2 += -rho * GradZ * d[M332];
    GwF1 += -rho * GradZ * d[M331];
    GradZ = -a.Grav * rho * Js / (d[M331] + d[M332]);
    Gu += GradZ * (d[M311] + d[M312]);
    Gv += GradZ * (d[M321] + d[M322]);
    f2 += (GuF1 + GuF2 - Gu / rho) / Ms;
    f3 += (GvF1 + GvF2 - Gv / rho) / Ms;
    if (has_hi) {
      GradZ = 0.5 * (p_p - p);
      const double Gradw = GradZ * (d[M332] + d331_p);
      const double Gradp = -(Gradw + a.Grav * (rho * d[MJ2] + rho_p * J1_p));
      f4up += (GwF2 + Gradp) / den_up;
    }
    if (has_lo) f4lo += GwF1 / den_lo;
  }
  __syncthreads();
  // ---- phase 2: second hyperviscosity apply, horizontal mass / theta flux ----
  const double c_own = rth / rho;
  if (in_tile) {
    T.at(0, node, l) = Curl;
    T.at(1, node, l) = Div;
    T.at(2, node, l) = ThCx;
    T.at(3, node, l) = ThCy;
    T.at(4, node, l) = (-rho) * (S11 * u + S12 * v);  // Contra12(-Rho, u, v)
    T.at(5, node, l) = (-rho) * (S21 * u + S22 * v);
    T.at(6, node, l) = c_own;
  }
  __syncthreads();
  double f1 = 0, f5 = 0;
  if (valid) {
    const double DxCurl = dx<N>(DWs, T, 0, I, J, l);
    const double DyCurl = dy<N>(DWs, T, 0, I, J, l);
    const double DxDiv = dx<N>(DWs, T, 1, I, J, l);
    const double DyDiv = dy<N>(DWs, T, 1, I, J, l);
    const double DivTh = dxy<N>(DWs, T, 2, 3, I, J, l);
    const double FuC = S12 * DxCurl + S22 * DyCurl;
    const double FvC = -S11 * DxCurl - S21 * DyCurl;
    const double FuD = S11 * DxDiv + S21 * DyDiv;
    const double FvD = S12 * DxDiv + S22 * DyDiv;
    f2 += -(-a.Kc * FuC + a.Kg * FuD) / Ms;
    f3 += -(-a.Kc * FvC + a.Kg * FvD) / Ms;
    f5 += -a.Kd * DivTh / Ms;
    // DivRhoTrUpwind3New2Kernel!: horizontal part
    const double DivRhoTr = dxy_prod<N>(Ds, T, 4, 5, 6, I, J, l);
    const double DivRho = dxy<N>(Ds, T, 4, 5, I, J, l);
    f5 += DivRhoTr / Ms;
    f1 += DivRho / Ms;
    // vertical part: upper interface iz and lower interface iz-1 (the
    // reference adds both from the owner of the interface)
    const double cm = rth_m / rho_m, cp = rth_p / rho_p;
    if (has_hi) {
      const double wCon = rho * (d[M312] * u + d[M322] * v + d[M332] * w) +
                          rho_p * (d311_p * u_p + d321_p * v_p + d331_p * w);
      const double cLL = has_lo ? cm : c_own, JLL = has_lo ? Jc_m : Js;
      double cFL, cFR;
      recu4(cLL, c_own, cp, rth_pp / rho_pp, JLL, Js, Jc_p, Jc_pp, cFL, cFR);
      const double Flux = 0.25 * ((fabs(wCon) + wCon) * cFL + (-fabs(wCon) + wCon) * cFR);
      f5 += -Flux / Ms;
      f1 += -(0.5 * wCon) / Ms;
    }
    if (has_lo) {
      const double wCon = rho_m * (d312_m * u_m + d322_m * v_m + d332_m * w_m) +
                          rho * (d[M311] * u + d[M321] * v + d[M331] * w_m);
      double cFL, cFR;
      const double cRR = has_hi ? cp : c_own, JRR = has_hi ? Jc_p : Js;
      recu4(rth_mm / rho_mm, cm, c_own, cRR, Jc_mm, Jc_m, Js, JRR, cFL, cFR);
      const double Flux = 0.25 * ((fabs(wCon) + wCon) * cFL + (-fabs(wCon) + wCon) * cFR);
      f5 += Flux / Ms;
      f1 += (0.5 * wCon) / Ms;
    }
  }
  __syncthreads();
  // ---- phase 3: vector-invariant momentum + Coriolis ----
  // U1 = 1/2 (tempuZ[2]-tempuZ[1]) + sum_k (D[I,k] tempuZ1(k,J) + D[J,k] tempuZ2(I,k))
  // and likewise V, W for the bottom (K=1) and top (K=2) vertical node; the
  // horizontal sums accumulate onto the vertical term in the reference's
  // order, six LDS tiles per round.
  const double tuZ1 = d[M331] * v - d[M321] * w_lo, tuZ2 = d[M332] * v - d[M322] * w_hi;
  const double tvZ1 = d[M331] * u - d[M311] * w_lo, tvZ2 = d[M332] * u - d[M312] * w_hi;
  const double twZ1 = d[M321] * u - d[M311] * v, twZ2 = d[M322] * u - d[M312] * v;
  double U1 = 0.5 * (tuZ2 - tuZ1), U2 = U1;
  double V1 = 0.5 * (tvZ2 - tvZ1), V2 = V1;
  double W1 = 0.5 * (twZ2 - twZ1), W2 = W1;
  if (in_tile) {
    T.at(0, node, l) = -d[M121] * w_lo;  // tempuZ1, K=1
    T.at(1, node, l) = -d[M221] * w_lo;  // tempuZ2
    T.at(2, node, l) = -d[M111] * w_lo;  // tempvZ1
    T.at(3, node, l) = -d[M211] * w_lo;  // tempvZ2
    T.at(4, node, l) = d[M121] * u - d[M111] * v;  // tempwZ1
    T.at(5, node, l) = d[M221] * u - d[M211] * v;  // tempwZ2
  }
  __syncthreads();
  if (valid) {
    U1 = dxy_acc<N>(U1, Ds, T, 0, 1, I, J, l);
    V1 = dxy_acc<N>(V1, Ds, T, 2, 3, I, J, l);
    W1 = dxy_acc<N>(W1, Ds, T, 4, 5, I, J, l);
  }
  __syncthreads();
  if (in_tile) {
    T.at(0, node, l) = -d[M122] * w_hi;  // K=2
    T.at(1, node, l) = -d[M222] * w_hi;
    T.at(2, node, l) = -d[M112] * w_hi;
    T.at(3, node, l) = -d[M212] * w_hi;
    T.at(4, node, l) = d[M122] * u - d[M112] * v;
    T.at(5, node, l) = d[M222] * u - d[M212] * v;
  }
  __syncthreads();
  if (valid) {
    U2 = dxy_acc<N>(U2, Ds, T, 0, 1, I, J, l);
    V2 = dxy_acc<N>(V2, Ds, T, 2, 3, I, J, l);
    W2 = dxy_acc<N>(W2, Ds, T, 4, 5, I, J, l);
    double FU = 0, FV = 0, FW1 = 0, FW2 = 0;
    if (has_lo) {
      const double tuZ2m = d332_m * v_m - d322_m * w_lo;
      const double tvZ2m = d332_m * u_m - d312_m * w_lo;
      const double twZ2m = d322_m * u_m - d312_m * v_m;
      const double FluxUZ = 0.5 * (tuZ1 - tuZ2m), FluxVZ = 0.5 * (tvZ1 - tvZ2m), FluxWZ = 0.5 * (twZ1 - twZ2m);
      FU += (-v * FluxWZ - w_lo * FluxVZ);
      FV += (u * FluxWZ - w_lo * FluxUZ);
      FW1 += (u * FluxVZ + v * FluxUZ);
    }
    if (has_hi) {
      const double tuZ1p = d331_p * v_p - d321_p * w_hi;
      const double tvZ1p = d331_p * u_p - d311_p * w_hi;
      const double twZ1p = d321_p * u_p - d311_p * v_p;
      const double FluxUZ = 0.5 * (tuZ1p - tuZ2), FluxVZ = 0.5 * (tvZ1p - tvZ2), FluxWZ = 0.5 * (twZ1p - twZ2);
      FU += (-v * FluxWZ - w_hi * FluxVZ);
      FV += (u * FluxWZ - w_hi * FluxUZ);
      FW2 += (u * FluxVZ + v * FluxUZ);
    }
    // CoriolisShallow: d[MCOR] = 2 Omega sin(lat) at the layer midpoint
    FU += (v * d[MCOR]) * Js;
    FV += (-u * d[MCOR]) * Js;
    f2 += (-v * W1 - w_lo * V1 - v * W2 - w_hi * V2 + FU) / Ms;
    f3 += (u * W1 - w_lo * U1 + u * W2 - w_hi * U2 + FV) / Ms;
    if (has_lo) f4lo += rho * (u * V1 + v * U1 + FW1) / den_lo;
    if (has_hi) f4up += rho * (u * V2 + v * U2 + FW2) / den_up;
  }
  __syncthreads();
  // ---- phase 4: F4 from the layer above (same element, same node) ----
  if (in_tile) T.at(0, node, l) = f4lo;
  __syncthreads();
  if (valid) {
    if (has_hi && l + 1 < nzt) f4up += T.at(0, node, l + 1);
    F[ui] += f1;
    F[ui + vs] += f2;
    F[ui + 2 * vs] += f3;
    if (has_hi) F[ui + 3 * vs] += f4up;
    F[ui + 4 * vs] += f5;
    // first level of a chunk above the first: its contribution to the
    // interface below belongs to the previous chunk's last level
    if (l == 0 && has_lo) F[ui - 1 + 3 * vs] += f4lo;
  }
}

// ------------------------------------------------------------------------
// PressureKernel! with Models.Dry (DiagnosticKernel.jl:96-115)
__global__ void k_prepare(KArgs a, const double *__restrict__ U, double *__restrict__ AuxG, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double rho = U[i], rth = U[i + 4 * n];
    const double p = eos_p(a, rth);
    const double T = p / (a.Rd * rho);
    AuxG[i] = p;
    AuxG[i + n] = T;
    AuxG[i + 2 * n] = pow(a.p0 / p, a.Rd / a.Cpd) * T;
  }
}

__global__ void k_fill(double *__restrict__ x, size_t n, double val) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = val;
}

// JacSchurKernel! (src/DyCore/JacSchur.jl:29-75); p = AuxG[:,:,4]
__global__ void k_jac(KArgs a, const double *__restrict__ U, const double *__restrict__ P4,
                      const double *__restrict__ dz, double *__restrict__ JRhoW, double *__restrict__ JWRho,
                      double *__restrict__ JWRhoTh, double *__restrict__ JRhoThW) {
  const int Nz = a.Nz, n = Nz - 1;
  const size_t total = (size_t)a.NumG * n;
  const size_t vs = (size_t)a.NumG * Nz;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const size_t ig = q / n;
    const int iz = (int)(q - ig * n);
    const size_t i0 = ig * Nz + iz;
    const double RhoL = U[i0], RhoR = U[i0 + 1];
    const double RhoThL = U[i0 + 4 * vs] + P4[i0], RhoThR = U[i0 + 1 + 4 * vs] + P4[i0 + 1];
    const double dzL = dz[i0], dzR = dz[i0 + 1];
    const double RhoF = (RhoL * dzL + RhoR * dzR) / (dzL + dzR);
    const size_t o = 2 * q;
    JRhoW[o] = -RhoF / dzL;
    JRhoW[o + 1] = RhoF / dzR;
    const double dPdThL = eos_dpdrhoth(a.Rd, a.p0, a.kappa, RhoThL);
    const double dPdThR = eos_dpdrhoth(a.Rd, a.p0, a.kappa, RhoThR);
    JWRhoTh[o] = -dPdThR / RhoF / (0.5 * (dzL + dzR));
    JWRhoTh[o + 1] = dPdThL / RhoF / (0.5 * (dzL + dzR));
    JWRho[o] = -a.Grav * dzR / RhoF / (dzL + dzR);
    JWRho[o + 1] = -a.Grav * dzL / RhoF / (dzL + dzR);
    const double RhoThF = (RhoThL * dzL + RhoThR * dzR) / (dzL + dzR);
    JRhoThW[o] = -RhoThF / dzL;
    JRhoThW[o + 1] = RhoThF / dzR;
  }
}

// SchurSolveGPU! (src/Integration/SchurSolve.jl:45-177): tri build,
// SchurSolveFKernel!, triSolve! (Thomas) and SchurSolveBKernel! for one
// column per thread.  bwork holds the Thomas right-hand side copy.
__global__ void k_schur(KArgs a, double *__restrict__ k, double *__restrict__ v, const double *__restrict__ JRhoW,
                        const double *__restrict__ JWRho, const double *__restrict__ JWRhoTh,
                        const double *__restrict__ JRhoThW, double *__restrict__ tri, double *__restrict__ bwork,
                        double fac, int comp_tri) {
  const int Nz = a.Nz, n = Nz - 1;
  const size_t ig = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= (size_t)a.NumG) return;
  const size_t vs = (size_t)a.NumG * Nz;
  const size_t c0 = ig * Nz;
  const size_t jb = ig * (size_t)n * 2, tb = ig * (size_t)n * 3;
#define JBk(A, c, i) A[jb + 2 * (size_t)(i) + (c)]
#define TRk(c, i) tri[tb + 3 * (size_t)(i) + (c)]
  if (comp_tri) {
    const double invfac = 1.0 / fac, invfac2 = invfac / fac;
    for (int i = 0; i < n; ++i) {
      TRk(1, i) = invfac2 - JBk(JWRho, 1, i) * JBk(JRhoW, 0, i) - JBk(JWRho, 0, i) * JBk(JRhoW, 1, i) -
                  JBk(JWRhoTh, 1, i) * JBk(JRhoThW, 0, i) - JBk(JWRhoTh, 0, i) * JBk(JRhoThW, 1, i);
      TRk(0, i) = (i == 0) ? 0.0
                           : -JBk(JWRho, 0, i - 1) * JBk(JRhoW, 0, i) - JBk(JWRhoTh, 0, i - 1) * JBk(JRhoThW, 0, i);
      TRk(2, i) = (i == n - 1) ? 0.0
                               : -JBk(JWRho, 1, i + 1) * JBk(JRhoW, 1, i) -
                                     JBk(JWRhoTh, 1, i + 1) * JBk(JRhoThW, 1, i);
    }
  }
  double *v1 = v + c0, *v4 = v + c0 + 3 * vs, *v5 = v + c0 + 4 * vs;
  double *k4 = k + c0 + 3 * vs;
  double *b = bwork + ig * (size_t)n;
  for (int iz = 0; iz < n; ++iz) {
    v4[iz] = v4[iz] / fac + JBk(JWRho, 1, iz) * v1[iz] + JBk(JWRho, 0, iz) * v1[iz + 1] +
             JBk(JWRhoTh, 1, iz) * v5[iz] + JBk(JWRhoTh, 0, iz) * v5[iz + 1];
    b[iz] = v4[iz];
  }
  k4[n] = 0.0;
  // triSolve!: x in k4
  k4[0] = TRk(1, 0);
  for (int i = 0; i < n - 1; ++i) {
    const double q = TRk(2, i) / k4[i];
    k4[i + 1] = TRk(1, i + 1) - q * TRk(0, i + 1);
    b[i + 1] = b[i + 1] - q * b[i];
  }
  k4[n - 1] = b[n - 1] / k4[n - 1];
  for (int i = n - 2; i >= 0; --i) k4[i] = (b[i] - TRk(0, i + 1) * k4[i + 1]) / k4[i];
  // SchurSolveBKernel!
  double *k1 = k + c0, *k2 = k + c0 + vs, *k3 = k + c0 + 2 * vs, *k5 = k + c0 + 4 * vs;
  const double *v2 = v + c0 + vs, *v3 = v + c0 + 2 * vs;
  for (int iz = 0; iz < Nz; ++iz) {
    if (iz == 0) {
      v1[iz] += JBk(JRhoW, 0, iz) * k4[iz];
      v5[iz] += JBk(JRhoThW, 0, iz) * k4[iz];
    } else if (iz == Nz - 1) {
      v1[iz] += JBk(JRhoW, 1, iz - 1) * k4[iz - 1];
      v5[iz] += JBk(JRhoThW, 1, iz - 1) * k4[iz - 1];
    } else {
      v1[iz] += JBk(JRhoW, 0, iz) * k4[iz] + JBk(JRhoW, 1, iz - 1) * k4[iz - 1];
      v5[iz] += JBk(JRhoThW, 0, iz) * k4[iz] + JBk(JRhoThW, 1, iz - 1) * k4[iz - 1];
    }
    k1[iz] = fac * v1[iz];
    k2[iz] = fac * v2[iz];
    k3[iz] = fac * v3[iz];
    k5[iz] = fac * v5[iz];
  }
#undef JBk
#undef TRk
}

// y = x0 + a1*x1 (+ a2*x2), evaluated as ((x0 + a1 x1) + a2 x2) like the
// successive broadcasts of RosenbrockSchur!.
__global__ void k_axpy(double *__restrict__ y, const double *__restrict__ x0, const double *__restrict__ x1,
                       double a1, const double *__restrict__ x2, double a2, int nterm, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double r = x0[i];
    if (nterm >= 1) r = r + a1 * x1[i];
    if (nterm >= 2) r = r + a2 * x2[i];
    y[i] = r;
  }
}

// y = x0 + a1 x1 + a2 x2 + a3 x3 (final RW update)
__global__ void k_axpy3(double *__restrict__ y, const double *__restrict__ x0, const double *__restrict__ x1,
                        double a1, const double *__restrict__ x2, double a2, const double *__restrict__ x3, double a3,
                        size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double r = x0[i];
    r = r + a1 * x1[i];
    r = r + a2 * x2[i];
    r = r + a3 * x3[i];
    y[i] = r;
  }
}

// ------------------------------------------------------------------------
// Context set-up: level-fastest metric for the element kernels.
// met[e][c][ID][iz] from dXdxI (3,3,2,NN,Nz,NF), J (NN,2,Nz,NF), X (NN,2,3,Nz,NF)
__global__ void k_relayout_metric(const double *__restrict__ dXdxI, const double *__restrict__ Jm,
                                  const double *__restrict__ X, double *__restrict__ met, double *__restrict__ met1,
                                  int NN, int Nz, int NF, double Omega) {
  const size_t total = (size_t)NF * NN * Nz;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(q % Nz);
    const size_t r = q / Nz;
    const int ID = (int)(r % NN);
    const size_t e = r / NN;
    const size_t pt = (size_t)ID + (size_t)NN * ((size_t)iz + (size_t)Nz * e);  // (ID, iz, e)
    const double *dp = dXdxI + 18 * pt;  // dXdxI[i,j,K] at dp[i + 3 j + 9 K]
#define DXI(i, j, K) dp[(i - 1) + 3 * (j - 1) + 9 * (K - 1)]
    const double vals[CGD_NMET - 1] = {DXI(1, 1, 1), DXI(1, 1, 2), DXI(1, 2, 1), DXI(1, 2, 2), DXI(2, 1, 1),
                                       DXI(2, 1, 2), DXI(2, 2, 1), DXI(2, 2, 2), DXI(3, 1, 1), DXI(3, 1, 2),
                                       DXI(3, 2, 1), DXI(3, 2, 2), DXI(3, 3, 1), DXI(3, 3, 2),
                                       Jm[(size_t)ID + (size_t)NN * (0 + 2 * ((size_t)iz + (size_t)Nz * e))],
                                       Jm[(size_t)ID + (size_t)NN * (1 + 2 * ((size_t)iz + (size_t)Nz * e))]};
    const size_t NNZ = (size_t)NN * Nz;
    const size_t mb = (e * CGD_NMET * NN + ID) * (size_t)Nz + iz;
    for (int c = 0; c < CGD_NMET - 1; ++c) met[mb + c * NNZ] = vals[c];
    // CoriolisShallow at the layer midpoint (Coriolis.jl:16-28, MomentumKernel.jl:156-159)
#define XA(K, c) X[(size_t)ID + (size_t)NN * ((K) + 2 * ((c) + 3 * ((size_t)iz + (size_t)Nz * e)))]
    const double x = 0.5 * (XA(0, 0) + XA(1, 0));
    const double y = 0.5 * (XA(0, 1) + XA(1, 1));
    const double z = 0.5 * (XA(0, 2) + XA(1, 2));
    const double sinlat = z / sqrt(x * x + y * y + z * z);
    met[mb + MCOR * NNZ] = 2.0 * Omega * sinlat;
    const size_t m1 = (e * CGD_NMET1 * NN + ID) * (size_t)Nz + iz;
    met1[m1] = DXI(1, 1, 1) + DXI(1, 1, 2);
    met1[m1 + NNZ] = DXI(1, 2, 1) + DXI(1, 2, 2);
    met1[m1 + 2 * NNZ] = DXI(2, 1, 1) + DXI(2, 1, 2);
    met1[m1 + 3 * NNZ] = DXI(2, 2, 1) + DXI(2, 2, 2);
    met1[m1 + 4 * NNZ] = vals[14] + vals[15];
#undef DXI
#undef XA
  }
}

__global__ void k_msum(const double *__restrict__ M, double *__restrict__ Ms, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    Ms[i] = M[i] + M[i + n];
}

template __global__ void k_fcn1<4>(KArgs, const int *, int, const double *, double *);
template __global__ void k_fcn1<5>(KArgs, const int *, int, const double *, double *);
template __global__ void k_fcn2<4>(KArgs, const int *, int, const double *, const double *, const double *, double *);
template __global__ void k_fcn2<5>(KArgs, const int *, int, const double *, const double *, const double *, double *);
