// kernels_solve.h -- device code of the supernodal triangular solves, the
// permutations, the FP64 residual SpMV and the refinement helpers (included
// once, by mxc_device.hip).
//
// Replaces CHOLMOD/Supernodal/t_cholmod_super_solve_worker.c (lsolve :20-303,
// ltsolve :309-580; the reference has no device solve) and the permutation
// steps of CHOLMOD/Cholesky/t_cholmod_psolve_worker.c:40-110.  Diagonal
// blocks are applied through their inverses (kept from the factorization),
// so every step is a small dense mat-vec; off-diagonal rows of a supernode are
// split across workgroups.  X is column-major n x nrhs (ld n), permuted order.
#pragma once
#include "mxc_kernels.h"

namespace mxc {

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// Y[k, r] = (T) B[Perm[k], r]
template <typename T>
__global__ void k_perm_gather(const double *__restrict__ B, const long long *__restrict__ Perm,
                              long long n, int nrhs, T *__restrict__ Y) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    long long pk = Perm[k];
    for (int r = 0; r < nrhs; r++) Y[k + r * n] = (T)B[pk + r * n];
}

// X[Perm[k], r] = Y[k, r]
template <typename T>
__global__ void k_perm_scatter(const T *__restrict__ Y, const long long *__restrict__ Perm,
                               long long n, int nrhs, double *__restrict__ X) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    long long pk = Perm[k];
    for (int r = 0; r < nrhs; r++) X[pk + r * n] = (double)Y[k + r * n];
}

// forward, diagonal block: x_b = inv(L_bb) x_b   (one wave per task)
template <typename T>
__global__ __launch_bounds__(64) void k_fwd_diag(const SolveTask *__restrict__ tasks,
                                                 const T *__restrict__ W, T *__restrict__ X,
                                                 long long n, int nrhs) {
    __shared__ T xs[NB];
    const SolveTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    T *xb = X + t.k1 + t.c0;
    for (int r = 0; r < nrhs; r++) {
        if (lane < t.nb) xs[lane] = xb[r * n + lane];
        __syncthreads();
        if (lane < t.nb) {
            T s = 0;
            const T *iv = W + t.winv + lane;
            for (int k = 0; k <= lane; k++) s += iv[(long long)k * t.nb] * xs[k];
            xb[r * n + lane] = s;
        }
        __syncthreads();
    }
}

// forward, rows [r0, r1) below block t:  x[g(r)] -= L[r, c0:c0+nb] x_b
// g(r) = k1 + r for the supernode's own columns (direct), Ls[psi + r] below
// them (atomic: sibling supernodes of one level share ancestor rows).
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_off(const SolveTask *__restrict__ tasks,
                                                 const T *__restrict__ W,
                                                 const int *__restrict__ Ls, T *__restrict__ X,
                                                 long long n, int nrhs) {
    __shared__ T xs[8][NB];
    const SolveTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const T *xb = X + t.k1 + t.c0;
    const int r = t.r0 + tid;
    const bool active = r < t.r1;
    long long g = 0;
    if (active) g = (r < t.nscol) ? (long long)(t.k1 + r) : (long long)Ls[t.psi + r];
    const T *Lcol = W + t.lx + (long long)t.c0 * t.nsrow + r;
    for (int r0 = 0; r0 < nrhs; r0 += 8) {
        const int nr = min(8, nrhs - r0);
        for (int e = tid; e < nr * NB; e += 256) {
            int rr = e / NB, k = e % NB;
            xs[rr][k] = (k < t.nb) ? xb[(r0 + rr) * n + k] : (T)0;
        }
        __syncthreads();
        if (active) {
            T acc[8];
#pragma unroll
            for (int q = 0; q < 8; q++) acc[q] = 0;
            for (int k = 0; k < t.nb; k++) {
                const T l = Lcol[(long long)k * t.nsrow];
#pragma unroll
                for (int q = 0; q < 8; q++) acc[q] += l * xs[q][k];
            }
            for (int q = 0; q < nr; q++) {
                if (r < t.nscol) X[g + (r0 + q) * n] -= acc[q];
                else atomic_add(X + g + (r0 + q) * n, -acc[q]);
            }
        }
        __syncthreads();
    }
}

// backward, rows [r0, r1) below block t:  x_b -= L[r0:r1, c0:c0+nb]' x[g(r0:r1)]
// 32-row slabs of L are staged through LDS (coalesced loads along rows), then
// thread (k, q) sums its column over the slab.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_off(const SolveTask *__restrict__ tasks,
                                                 const T *__restrict__ W,
                                                 const int *__restrict__ Ls, T *__restrict__ X,
                                                 long long n, int nrhs) {
    constexpr int SLAB = 32;
    __shared__ T Lt[NB][SLAB + 1];
    __shared__ T xr[8][SLAB];
    __shared__ T part[4][8][NB];
    const SolveTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const int k = tid & (NB - 1), q = tid >> 6;
    T *xb = X + t.k1 + t.c0;
    for (int rb = 0; rb < nrhs; rb += 8) {
        const int nr = min(8, nrhs - rb);
        T acc[8];
#pragma unroll
        for (int z = 0; z < 8; z++) acc[z] = 0;
        for (int s0 = t.r0; s0 < t.r1; s0 += SLAB) {
            // stage L[s0:s0+SLAB, c0:c0+nb] and x[g(rows)]
            for (int e = tid; e < SLAB * NB; e += 256) {
                int rr = e % SLAB, kk = e / SLAB;
                int r = s0 + rr;
                T v = 0;
                if (r < t.r1 && kk < t.nb) v = W[t.lx + (long long)(t.c0 + kk) * t.nsrow + r];
                Lt[kk][rr] = v;
            }
            for (int e = tid; e < nr * SLAB; e += 256) {
                int z = e / SLAB, rr = e % SLAB;
                int r = s0 + rr;
                T v = 0;
                if (r < t.r1) {
                    long long g = (r < t.nscol) ? (long long)(t.k1 + r) : (long long)Ls[t.psi + r];
                    v = X[g + (rb + z) * n];
                }
                xr[z][rr] = v;
            }
            __syncthreads();
            for (int rr = q; rr < SLAB; rr += 4) {
                const T l = Lt[k][rr];
#pragma unroll
                for (int z = 0; z < 8; z++) acc[z] += l * xr[z][rr];
            }
            __syncthreads();
        }
#pragma unroll
        for (int z = 0; z < 8; z++) part[q][z][k] = acc[z];
        __syncthreads();
        if (q == 0 && k < t.nb) {
            for (int z = 0; z < nr; z++) {
                T s = part[0][z][k] + part[1][z][k] + part[2][z][k] + part[3][z][k];
                atomic_add(xb + (rb + z) * n + k, -s);
            }
        }
        __syncthreads();
    }
}

// backward, diagonal block: x_b = inv(L_bb)' x_b
template <typename T>
__global__ __launch_bounds__(64) void k_bwd_diag(const SolveTask *__restrict__ tasks,
                                                 const T *__restrict__ W, T *__restrict__ X,
                                                 long long n, int nrhs) {
    __shared__ T xs[NB];
    const SolveTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    T *xb = X + t.k1 + t.c0;
    for (int r = 0; r < nrhs; r++) {
        if (lane < t.nb) xs[lane] = xb[r * n + lane];
        __syncthreads();
        if (lane < t.nb) {
            T s = 0;
            const T *iv = W + t.winv + (long long)lane * t.nb;   // column `lane` of inv
            for (int k = lane; k < t.nb; k++) s += iv[k] * xs[k];
            xb[r * n + lane] = s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FP64 residual R = B - A X with A held as full CSR (both triangles), values
// gathered from tril(A) by src.  One wave per 64 rows, one thread per row.
// ---------------------------------------------------------------------------
__global__ void k_gather_vals(const double *__restrict__ Ax, const long long *__restrict__ src,
                              long long nnz, double *__restrict__ vals) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; p < nnz; p += stride) vals[p] = Ax[src[p]];
}

__global__ void k_residual(const long long *__restrict__ rowptr, const int *__restrict__ col,
                           const double *__restrict__ vals, const double *__restrict__ B,
                           const double *__restrict__ X, double *__restrict__ R, long long n,
                           int nrhs) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long p0 = rowptr[i], p1 = rowptr[i + 1];
    for (int r = 0; r < nrhs; r++) {
        const double *x = X + r * n;
        double s = 0;
        for (long long p = p0; p < p1; p++) s += vals[p] * x[col[p]];
        R[i + r * n] = B[i + r * n] - s;
    }
}

// out[r] += sum_i v[i, r]^2  (out zeroed by the caller)
__global__ void k_sumsq(const double *__restrict__ V, long long n, int nrhs,
                        double *__restrict__ out) {
    __shared__ double red[256];
    for (int r = 0; r < nrhs; r++) {
        double s = 0;
        for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (long long)gridDim.x * blockDim.x) {
            double v = V[i + r * n];
            s += v * v;
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomic_add(out + r, red[0]);
        __syncthreads();
    }
}

// out[g * nrhs + r] = sum over rows of segment g of v[i, r]^2; one block per segment
__global__ void k_seg_sumsq(const double *__restrict__ V, long long n, int nrhs,
                            const long long *__restrict__ segoff, double *__restrict__ out) {
    __shared__ double red[256];
    const int g = blockIdx.x;
    const long long a = segoff[g], b = segoff[g + 1];
    for (int r = 0; r < nrhs; r++) {
        double s = 0;
        for (long long i = a + threadIdx.x; i < b; i += blockDim.x) {
            double v = V[i + r * n];
            s += v * v;
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(long long)g * nrhs + r] = red[0];
        __syncthreads();
    }
}

// X += D (converted), over n*nrhs entries
template <typename T>
__global__ void k_axpy_update(double *__restrict__ X, const T *__restrict__ D, long long len) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < len; i += stride) X[i] += (double)D[i];
}

}  // namespace mxc
