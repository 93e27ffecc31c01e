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

// Stage the nb x nb inverse of block t (ld nb) into LDS Iv[64][65] with every
// thread's loads in flight together (256 threads, 16 each).
template <typename T>
__device__ __forceinline__ void stage_inverse(const T *__restrict__ W, long long winv, int nb,
                                              T (*Iv)[NB + 1], int tid) {
    const int i = tid & (NB - 1), c0 = tid >> 6;
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int c = c0 + 4 * q;
        v[q] = (i < nb && c < nb) ? W[winv + i + (long long)c * nb] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) Iv[i][c0 + 4 * q] = v[q];
}

// forward, diagonal block: x_b = inv(L_bb) x_b
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_diag(const SolveTask *__restrict__ tasks,
                                                  const T *__restrict__ W, T *__restrict__ X,
                                                  long long n, int nrhs) {
    __shared__ T Iv[NB][NB + 1];
    __shared__ T xs[4][NB];
    const SolveTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    T *xb = X + t.k1 + t.c0;
    stage_inverse(W, t.winv, t.nb, Iv, tid);
    for (int r0 = 0; r0 < nrhs; r0 += 4) {
        const int r = r0 + q;
        if (r < nrhs) xs[q][lane] = (lane < t.nb) ? xb[r * n + lane] : (T)0;
        __syncthreads();
        if (r < nrhs && lane < t.nb) {
            T s = 0;
            for (int k = 0; k <= lane; k++) s += Iv[lane][k] * xs[q][k];
            xb[r * n + lane] = s;
        }
        __syncthreads();
    }
}

// forward, rows [r0, r1) below block t:  x[g(r)] -= L[r, c0:c0+nb] x_b
// g(r) = k1 + r for the supernode's own columns (direct), Ls[psi + r] below
// them (atomic: sibling supernodes of one level share ancestor rows).
// One thread per row; the row's 16-column chunks are loaded together.
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
    const T *Lrow = W + t.lx + (long long)t.c0 * t.nsrow + r;
    for (int rb = 0; rb < nrhs; rb += 8) {
        const int nr = min(8, nrhs - rb);
        for (int e = tid; e < 8 * NB; e += 256) {
            const int z = e / NB, k = e % NB;
            xs[z][k] = (z < nr && k < t.nb) ? xb[(rb + z) * n + k] : (T)0;
        }
        __syncthreads();
        if (active) {
            T acc[8];
#pragma unroll
            for (int z = 0; z < 8; z++) acc[z] = 0;
            for (int kc = 0; kc < t.nb; kc += 16) {
                T l[16];
#pragma unroll
                for (int k = 0; k < 16; k++)
                    l[k] = (kc + k < t.nb) ? Lrow[(long long)(kc + k) * t.nsrow] : (T)0;
#pragma unroll
                for (int k = 0; k < 16; k++)
#pragma unroll
                    for (int z = 0; z < 8; z++) acc[z] += l[k] * xs[z][kc + k];
            }
            for (int z = 0; z < nr; z++) {
                if (r < t.nscol) X[g + (rb + z) * n] -= acc[z];
                else atomic_add(X + g + (rb + z) * n, -acc[z]);
            }
        }
        __syncthreads();
    }
}

// Sum of 16 per-lane values over the 64 lanes of a wave, transposed: on
// return v[0] in lane l is the total of index (l >> 2) & 15 (every lane of
// a group of 4 holds the same total).  Halving butterfly: 17 shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum16(T (&v)[16], int lane) {
#pragma unroll
    for (int lvl = 0; lvl < 4; lvl++) {
        const int half = 8 >> lvl;            // 8, 4, 2, 1
        const int mask = 32 >> lvl;           // 32, 16, 8, 4
        const bool up = (lane & mask) != 0;
#pragma unroll
        for (int i = 0; i < half; i++) {
            const T keep = up ? v[i + half] : v[i];
            const T send = up ? v[i] : v[i + half];
            v[i] = keep + __shfl_xor(send, mask);
        }
    }
    T s = v[0];
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 1);
    return s;
}

// backward, rows [r0, r1) below block t:  x_b -= L[r0:r1, c0:c0+nb]' x[g(r0:r1)]
// One thread per row (coalesced reads of L along rows), per-wave transposed
// butterfly reduction over 16-column chunks, atomic add into x_b.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_off(const SolveTask *__restrict__ tasks,
                                                 const T *__restrict__ W,
                                                 const int *__restrict__ Ls, T *__restrict__ X,
                                                 long long n, int nrhs) {
    const SolveTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63;
    T *xb = X + t.k1 + t.c0;
    const int r = t.r0 + tid;
    const bool active = r < t.r1;
    long long g = 0;
    if (active) g = (r < t.nscol) ? (long long)(t.k1 + r) : (long long)Ls[t.psi + r];
    const T *Lrow = W + t.lx + (long long)t.c0 * t.nsrow + r;
    for (int kc = 0; kc < t.nb; kc += 16) {
        T l[16];
#pragma unroll
        for (int k = 0; k < 16; k++)
            l[k] = (active && kc + k < t.nb) ? Lrow[(long long)(kc + k) * t.nsrow] : (T)0;
        for (int z = 0; z < nrhs; z++) {
            const T xr = active ? X[g + z * n] : (T)0;
            T v[16];
#pragma unroll
            for (int k = 0; k < 16; k++) v[k] = l[k] * xr;
            const T s = wave_sum16(v, lane);
            const int k = kc + ((lane >> 2) & 15);
            if ((lane & 3) == 0 && k < t.nb) atomic_add(xb + z * n + k, -s);
        }
    }
}

// backward, diagonal block: x_b = inv(L_bb)' x_b
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_diag(const SolveTask *__restrict__ tasks,
                                                  const T *__restrict__ W, T *__restrict__ X,
                                                  long long n, int nrhs) {
    __shared__ T Iv[NB][NB + 1];
    __shared__ T xs[4][NB];
    const SolveTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    T *xb = X + t.k1 + t.c0;
    stage_inverse(W, t.winv, t.nb, Iv, tid);
    for (int r0 = 0; r0 < nrhs; r0 += 4) {
        const int r = r0 + q;
        if (r < nrhs) xs[q][lane] = (lane < t.nb) ? xb[r * n + lane] : (T)0;
        __syncthreads();
        if (r < nrhs && lane < t.nb) {
            T s = 0;
            for (int k = lane; k < t.nb; k++) s += Iv[k][lane] * xs[q][k];
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
