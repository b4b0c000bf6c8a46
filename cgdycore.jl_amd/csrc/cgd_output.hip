// Output on the device: interpolation of global-node fields to the cell
// centres of an OrdPrint x OrdPrint sub-grid of every element and layer
// (InterpolateKernel!, InterpolateRhoKernel!, InterpolateWBKernel!,
// src/Outputs/vtkOutputKernel.jl:49-158), written in the reference's
// cCell (OrdPrint, OrdPrint, Nz, NF) layout.  One thread per output cell;
// the sum over the element's nodes runs in the reference's (jP, iP) order.
#include "../../include/cgdycore_hip.h"

#include <hip/hip_runtime.h>

#include <string>

namespace {

thread_local std::string g_out_err;

int out_fail(int code, const std::string &msg) {
  g_out_err = msg;
  return code;
}

int blocks_for(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// mode 0: c; mode 1: c / rho; mode 2: the cell-centre w, 1/2 (w[Iz] + w[Iz-1])
// with the free-slip w0 = -(u dXdxI[3,1,1] + v dXdxI[3,2,1]) / dXdxI[3,3,1]
// below the first layer
__global__ void k_interpolate(int mode, int N, int Nz, int NF, int NumG, int P, const int64_t *__restrict__ Glob,
                              const double *__restrict__ Inter, const double *__restrict__ c,
                              const double *__restrict__ rho, const double *__restrict__ u,
                              const double *__restrict__ v, const double *__restrict__ dXdxI,
                              double *__restrict__ cCell) {
  const int NN = N * N;
  const size_t total = (size_t)P * P * Nz * NF;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const int I = (int)(q % P);
    const int J = (int)((q / P) % P);
    const size_t r = q / ((size_t)P * P);
    const int Iz = (int)(r % Nz);
    const size_t IF = r / Nz;
    double s = 0.0;
    int iD = 0;
    for (int jP = 0; jP < N; ++jP)
      for (int iP = 0; iP < N; ++iP, ++iD) {
        const size_t ind = (size_t)(Glob[iD + (size_t)NN * IF] - 1);
        // Inter[I,J,iP,jP] (column-major (P,P,N,N))
        const double L = Inter[I + (size_t)P * (J + (size_t)P * (iP + (size_t)N * jP))];
        const size_t o = (size_t)Iz + (size_t)Nz * ind;
        if (mode == 0) {
          s += L * c[o];
        } else if (mode == 1) {
          s += L * c[o] / rho[o];
        } else {
          double wb;
          if (Iz == 0) {
            const double *d = dXdxI + 18 * ((size_t)iD + (size_t)NN * ((size_t)Nz * IF));  // (.,.,K=1,iD,1,IF)
            wb = -(u[o] * d[2 + 3 * 0] + v[o] * d[2 + 3 * 1]) / d[2 + 3 * 2];
          } else {
            wb = c[o - 1];
          }
          s += 0.5 * L * (c[o] + wb);
        }
      }
    cCell[q] = s;
  }
}

}  // namespace

extern "C" const char *cgd_output_error(void) { return g_out_err.c_str(); }

extern "C" int cgd_interpolate(int mode, int N, int Nz, int NF, int NumG, int OrdPrint, const int64_t *Glob,
                               const double *Inter, const double *c, const double *rho, const double *u,
                               const double *v, const double *dXdxI, double *cCell, void *stream) {
  if (mode < 0 || mode > 2) return out_fail(CGD_ERR_ARG, "mode must be 0 (c), 1 (c/rho) or 2 (cell-centre w)");
  if (N < 2 || N > 8 || Nz < 1 || NF < 1 || NumG < 1 || OrdPrint < 1 || OrdPrint > 16)
    return out_fail(CGD_ERR_ARG, "bad dimension");
  if (!Glob || !Inter || !c || !cCell) return out_fail(CGD_ERR_ARG, "null array");
  if (mode == 1 && !rho) return out_fail(CGD_ERR_ARG, "mode 1 needs rho");
  if (mode == 2 && (!u || !v || !dXdxI)) return out_fail(CGD_ERR_ARG, "mode 2 needs u, v and dXdxI");
  const size_t total = (size_t)OrdPrint * OrdPrint * Nz * NF;
  hipLaunchKernelGGL(k_interpolate, dim3(blocks_for(total, 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), mode, N, Nz, NF, NumG, OrdPrint, Glob, Inter, c, rho, u,
                     v, dXdxI, cCell);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return out_fail(CGD_ERR_HIP, std::string("k_interpolate: ") + hipGetErrorString(e));
  return CGD_OK;
}
