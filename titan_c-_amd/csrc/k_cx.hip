// k_cx.hip -- exact FP64 charge-exchange source kernels and the reference's closed-form
// (Calc_sourse_MF_Bera) source (FP64, no contraction).
#define TITAN_FLUX_ONLY 1
#include "../../include/titan_gpu.h"
#include "bera.h"
#include "cx_device.h"
#include "launch.h"
#include "step_kernels.h"

namespace titan {

__global__ void __launch_bounds__(256) k_cx_exact(int64_t n, const double *__restrict__ in10, int nodes,
                                                  double *__restrict__ out5) {
    extern __shared__ double sc[];
    cx_fill_angles(sc, nodes);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *a = in10 + 10 * i;
    const double vp[3] = {a[2], a[3], a[4]}, vh[3] = {a[7], a[8], a[9]};
    double o[5];
    cx_source_one(a[0], a[1], vp, a[5], a[6], vh, nodes, nodes, sc, o);
#pragma unroll
    for (int k = 0; k < 5; k++) out5[5 * i + k] = o[k];
}

// Per cell: n_p = rho, c_p^2 = p/rho, V_p from the state; neutrals [5] = n_H, c_H^2, V_H.
__device__ __forceinline__ void cx_cell(const double *__restrict__ prim, const double *__restrict__ neutrals,
                                        int64_t c, int nodes, const double *sc, double o[5]) {
    const double *s = prim + 8 * c;
    const double *h = neutrals + 5 * c;
    const double vp[3] = {s[1], s[2], s[3]}, vh[3] = {h[2], h[3], h[4]};
    cx_source_one(s[0], s[4] / s[0], vp, h[0], h[1], vh, nodes, nodes, sc, o);
}

__global__ void __launch_bounds__(256) k_cx_cells(int64_t nc, const double *__restrict__ prim,
                                                  const double *__restrict__ neutrals, int nodes,
                                                  double *__restrict__ src) {
    extern __shared__ double sc[];
    cx_fill_angles(sc, nodes);
    __syncthreads();
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double o[5];
    cx_cell(prim, neutrals, c, nodes, sc, o);
#pragma unroll
    for (int k = 0; k < 5; k++) src[5 * c + k] = o[k];
}

// Exact source for the cells of a list (the surrogate's out-of-range fallback).
__global__ void __launch_bounds__(256) k_cx_list(const int32_t *__restrict__ list, const unsigned long long *count,
                                                 int64_t cap, const double *__restrict__ prim,
                                                 const double *__restrict__ neutrals, int nodes,
                                                 double *__restrict__ src) {
    extern __shared__ double sc[];
    cx_fill_angles(sc, nodes);
    __syncthreads();
    const int64_t n = (int64_t)min((unsigned long long)cap, *count);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = list[i];
        double o[5];
        cx_cell(prim, neutrals, c, nodes, sc, o);
#pragma unroll
        for (int k = 0; k < 5; k++) src[5 * c + k] = o[k];
    }
}

// Cells that the surrogate step kernel left outside the training box (list and stash written by
// k_step_tile<*, true>, count in S->n_fallback) get the exact FP64 source and their update redone
// from the stashed flux sum and volumes; prim1 is the step's unchanged input layer.  The first
// kFixPar listed cells spread their quadrature over kFixSlices workgroups each (k_cx_fix_part:
// one slice of the N1 x N2 points per workgroup, thread-strided, fixed-order tree reduction into
// part), and k_cx_fix_finish sums the slices in order and updates them, one thread per cell; the
// cells past kFixPar take one workgroup each (k_cx_fixup from `first`).
__global__ void __launch_bounds__(256) k_cx_fix_part(const double *__restrict__ prim1,
                                                     const double *__restrict__ neutrals, int nodes,
                                                     const int32_t *__restrict__ list, int64_t cap,
                                                     const StepScalars *S, double *__restrict__ part) {
    extern __shared__ double sm[];
    double *sc = sm, *red = sm + 2 * nodes;
    const int64_t n = min(min((int64_t)S->n_fallback, cap), (int64_t)kFixPar);
    const int64_t items = n * kFixSlices;
    if ((int64_t)blockIdx.x >= items) return;  // whole workgroup
    cx_fill_angles(sc, nodes);
    __syncthreads();
    const int np = nodes * nodes;
    const int t = threadIdx.x, nt = blockDim.x;
    for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const int64_t i = it / kFixSlices;
        const int sl = (int)(it - i * kFixSlices);
        const int64_t c = list[i];
        const double *s = prim1 + 8 * c;
        const double *h = neutrals + 5 * c;
        const double vp[3] = {s[1], s[2], s[3]}, vh[3] = {h[2], h[3], h[4]};
        const CxFrame f = cx_frame(s[4] / s[0], vp, h[1], vh, nodes, nodes);
        const int p0 = (int)((int64_t)np * sl / kFixSlices), p1 = (int)((int64_t)np * (sl + 1) / kFixSlices);
        double J0 = 0.0, J1 = 0.0, J2 = 0.0;
        for (int p = p0 + t; p < p1; p += nt) {
            const int ii = p / nodes, jj = p - ii * nodes;
            double R0;
            const double wr = cx_radial(f, ii, R0);
            double gs, gpar, g2;
            cx_point(f, R0, jj, sc, gs, gpar, g2);
            const double w = wr * gs;
            J0 += w;
            J1 += w * gpar;
            J2 += w * g2;
        }
        red[t] = J0; red[nt + t] = J1; red[2 * nt + t] = J2;
        __syncthreads();
        for (int hh = nt / 2; hh > 0; hh >>= 1) {
            if (t < hh) {
                red[t] += red[t + hh];
                red[nt + t] += red[nt + t + hh];
                red[2 * nt + t] += red[2 * nt + t + hh];
            }
            __syncthreads();
        }
        if (t == 0) {
            part[3 * it] = red[0];
            part[3 * it + 1] = red[nt];
            part[3 * it + 2] = red[2 * nt];
        }
        __syncthreads();  // red is reused by the next item
    }
}

__global__ void __launch_bounds__(64) k_cx_fix_finish(DevMesh M, const double *__restrict__ prim1,
                                                      const double *__restrict__ neutrals, int nodes,
                                                      const int32_t *__restrict__ list,
                                                      const double *__restrict__ stash, int64_t cap, StepScalars *S,
                                                      const double *__restrict__ part, double *__restrict__ prim2) {
    const int64_t n = min(min((int64_t)S->n_fallback, cap), (int64_t)kFixPar);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = list[i];
    const double *s = prim1 + 8 * c;
    const double *h = neutrals + 5 * c;
    const double vp[3] = {s[1], s[2], s[3]}, vh[3] = {h[2], h[3], h[4]};
    const double cp2 = s[4] / s[0];
    const CxFrame f = cx_frame(cp2, vp, h[1], vh, nodes, nodes);
    double J0 = 0.0, J1 = 0.0, J2 = 0.0;
    for (int sl = 0; sl < kFixSlices; sl++) {
        const double *q = part + 3 * (i * kFixSlices + sl);
        J0 += q[0];
        J1 += q[1];
        J2 += q[2];
    }
    double o[5];
    cx_finish(f, J0, J1, J2, s[0], cp2, vp, h[0], h[1], vh, o);
    const double *w = stash + 11 * i;
    double Pf[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Pf[k] = w[k];
    cell_update_one(M, M.cell_code[c], c, prim1, w[9], w[10], Pf, o, S->dt, S, prim2);
}

__global__ void __launch_bounds__(256) k_cx_fixup(DevMesh M, const double *__restrict__ prim1,
                                                  const double *__restrict__ neutrals, int nodes,
                                                  const int32_t *__restrict__ list, const double *__restrict__ stash,
                                                  int64_t cap, int64_t first, StepScalars *S,
                                                  double *__restrict__ prim2) {
    extern __shared__ double sm[];
    double *sc = sm, *red = sm + 2 * nodes;
    const int64_t n = min((int64_t)S->n_fallback, cap);
    if (first + (int64_t)blockIdx.x >= n) return;  // whole workgroup
    cx_fill_angles(sc, nodes);
    __syncthreads();
    const double dt = S->dt;
    for (int64_t i = first + blockIdx.x; i < n; i += gridDim.x) {
        const int64_t c = list[i];
        const double *s = prim1 + 8 * c;
        const double *h = neutrals + 5 * c;
        const double vp[3] = {s[1], s[2], s[3]}, vh[3] = {h[2], h[3], h[4]};
        double o[5];
        cx_source_block(s[0], s[4] / s[0], vp, h[0], h[1], vh, nodes, nodes, sc, red, o);
        if (threadIdx.x == 0) {
            const double *w = stash + 11 * i;
            double Pf[9];
#pragma unroll
            for (int k = 0; k < 9; k++) Pf[k] = w[k];
            cell_update_one(M, M.cell_code[c], c, prim1, w[9], w[10], Pf, o, dt, S, prim2);
        }
        __syncthreads();
    }
}

void launch_cx_fixup(hipStream_t s, const DevMesh &M, const double *prim1, const double *neutrals, int nodes,
                     const int32_t *list, const double *stash, int64_t cap, double *part, StepScalars *S,
                     double *prim2) {
    const size_t lds = sizeof(double) * (2 * (size_t)nodes + 3 * 256);
    hipLaunchKernelGGL(k_cx_fix_part, dim3(kFixPar * kFixSlices / 4), dim3(256), lds, s, prim1, neutrals, nodes, list,
                       cap, S, part);
    hipLaunchKernelGGL(k_cx_fix_finish, dim3(kFixPar / 64), dim3(64), 0, s, M, prim1, neutrals, nodes, list, stash,
                       cap, S, (const double *)part, prim2);
    hipLaunchKernelGGL(k_cx_fixup, dim3(256), dim3(256), lds, s, M, prim1, neutrals, nodes, list, stash, cap,
                       (int64_t)kFixPar, S, prim2);
}

void launch_cx_exact(hipStream_t s, int64_t n, const double *in10, int nodes, double *out5) {
    hipLaunchKernelGGL(k_cx_exact, dim3((unsigned)((n + 255) / 256)), dim3(256), 16 * nodes, s, n, in10, nodes, out5);
}

void launch_cx_cells(hipStream_t s, int64_t nc, const double *prim, const double *neutrals, int nodes, double *src) {
    hipLaunchKernelGGL(k_cx_cells, dim3((unsigned)((nc + 255) / 256)), dim3(256), 16 * nodes, s, nc, prim, neutrals,
                       nodes, src);
}

// Calc_sourse_MF_Bera per cell: hfluid [Nc * n_h * 5] (rho_H, p_H, V_H per fluid); scal [Nc*2]
// (Q, rho_He; nullable: both 0) and the cell's zone code select the proton split (bera_zone).
__global__ void __launch_bounds__(256) k_bera_cells(int64_t nc, const double *__restrict__ prim,
                                                    const double *__restrict__ scal,
                                                    const uint32_t *__restrict__ cell_code,
                                                    const double *__restrict__ hfluid, int n_h, double a2,
                                                    double ddp, double *__restrict__ src) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double s[8];
    load_state(prim, c, s);
    const double Q = scal ? scal[2 * c] : 0.0, rho_He = scal ? scal[2 * c + 1] : 0.0;
    const int zone = bera_zone(Q, s[0], (int)((cell_code[c] >> 8) & 15u));
    double o[5];
    bera_source(s, rho_He, zone, n_h, hfluid + (int64_t)5 * n_h * c, a2, ddp, o);
#pragma unroll
    for (int k = 0; k < 5; k++) src[5 * c + k] = o[k];
}

static_assert(kBeraMaxH == TITAN_MAX_H_FLUIDS, "the hydrogen fluids the sources take");

// The hydrogen fluids' sources of each cell (with its plasma sources): hout [nc][n_h][5].
__global__ void __launch_bounds__(256) k_bera_cells_h(int64_t nc, const double *__restrict__ prim,
                                                      const double *__restrict__ scal,
                                                      const uint32_t *__restrict__ cell_code,
                                                      const double *__restrict__ hfluid, int n_h, double a2,
                                                      double ddp, double n_h_lism, const int8_t *__restrict__ arise,
                                                      double *__restrict__ src, double *__restrict__ hout) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double s[8];
    load_state(prim, c, s);
    const double Q = scal ? scal[2 * c] : 0.0, rho_He = scal ? scal[2 * c + 1] : 0.0;
    const int zone = bera_zone(Q, s[0], (int)((cell_code[c] >> 8) & 15u));
    double o[5];
    bera_source<true>(s, rho_He, zone, n_h, hfluid + (int64_t)5 * n_h * c, a2, ddp, o, hout + (int64_t)5 * n_h * c,
                      arise, n_h_lism);
#pragma unroll
    for (int k = 0; k < 5; k++) src[5 * c + k] = o[k];
}

void launch_bera_cells_h(hipStream_t s, int64_t nc, const double *prim, const double *scal, const uint32_t *cell_code,
                         const double *hfluid, int n_h, double a2, double ddp, double n_h_lism, const int8_t *arise,
                         double *src, double *hout) {
    hipLaunchKernelGGL(k_bera_cells_h, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, nc, prim, scal, cell_code,
                       hfluid, n_h, a2, ddp, n_h_lism, arise, src, hout);
}

void launch_bera_cells(hipStream_t s, int64_t nc, const double *prim, const double *scal, const uint32_t *cell_code,
                       const double *hfluid, int n_h, double a2, double ddp, double *src) {
    hipLaunchKernelGGL(k_bera_cells, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, nc, prim, scal, cell_code,
                       hfluid, n_h, a2, ddp, src);
}

void launch_cx_list(hipStream_t s, const int32_t *list, const unsigned long long *count, int64_t cap,
                    const double *prim, const double *neutrals, int nodes, double *src) {
    hipLaunchKernelGGL(k_cx_list, dim3(1024), dim3(256), 16 * nodes, s, list, count, cap, prim, neutrals, nodes, src);
}

}  // namespace titan
