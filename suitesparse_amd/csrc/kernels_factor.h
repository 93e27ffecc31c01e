// kernels_factor.h -- device code of the numeric factorization (included once,
// by mxc_device.hip).
//
// Replaces, in one device-resident pass over a level schedule of the
// supernodal elimination tree:
//   * zero + scatter of A into the supernodes  (t_cholmod_super_numeric_worker.c:360-520)
//   * descendant SYRK/GEMM + relative-map scatter-subtract
//       (worker :769-930; GPU: t_cholmod_gpu.c:381-617, kernelAddUpdate
//        cholmod_gpu_kernels.cu:43-54)
//   * diagonal POTRF and panel TRSM (worker :1023-1027, :1175-1181;
//       GPU: t_cholmod_gpu.c:784-1135, :1151-1393)
#pragma once
#include "mxc_kernels.h"

namespace mxc {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaOp;
// Accumulator layout of the 16x16x4 MFMAs: lane l, register v holds
// D[row(l, v)][l % 16].  FP64 interleaves rows by 4 (row = l/16 + 4v); FP32
// keeps each lane's 4 rows contiguous (row = 4(l/16) + v).
template <> struct MfmaOp<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MfmaOp<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ __forceinline__ void atomic_sub(double *p, double v) { unsafeAtomicAdd(p, -v); }
__device__ __forceinline__ void atomic_sub(float *p, float v) { unsafeAtomicAdd(p, -v); }

// ---------------------------------------------------------------------------
// Assembly of A into L: Lx = 0; Lx[pos[p]] = Ax[p]; diagonal += beta.
// pos folds the fill-reducing permutation, so the values of tril(A) go in in
// their original order (no host-side P A P').
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *__restrict__ p, long long n, T v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
__global__ void k_scatter_a(const double *__restrict__ Ax, const long long *__restrict__ pos,
                            long long nnz, T *__restrict__ W) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; p < nnz; p += stride) W[pos[p]] = (T)Ax[p];
}

// beta is read from device memory so that a captured factorization graph
// serves every beta
template <typename T>
__global__ void k_add_diag(const long long *__restrict__ dpos, long long n,
                           const double *__restrict__ beta, T *__restrict__ W) {
    const double b = *beta;
    if (b == 0.0) return;
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) W[dpos[k]] += (T)b;
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// ---------------------------------------------------------------------------
// Dense TS x TS update tile on MFMA (see UpdDesc), TS = 64 (4 waves, 256
// threads) or 32 (1 wave, 64 threads, for the many small (d,s) pairs).  Each
// wave owns a 32x32 block of C as 2x2 16x16 MFMA blocks.  The MFMA is fed X
// (C-column side) as its A operand and Y (C-row side) as its B operand, so
// each lane's 4 accumulators hold 4 C columns of one C row and 16
// consecutive lanes hold 16 consecutive C rows: stores into the
// column-major target are 16-element contiguous runs.
// ---------------------------------------------------------------------------
template <typename T, int MODE, int TS>
__global__ __launch_bounds__(TS * TS / 16) void k_tile(const UpdDesc *__restrict__ descs,
                                                      const Tile *__restrict__ tiles, T *__restrict__ W,
                                                      const int *__restrict__ Ls) {
    typedef typename MfmaOp<T>::acc_t acc_t;
    constexpr int NT = TS * TS / 16;            // threads: 64 per 32x32 wave block
    constexpr int KG = NT / TS;                 // k groups staged per pass
    constexpr int PER = (BK * TS) / NT;         // elements of X (and of Y) each thread stages
    constexpr int WPR = TS / 32;                // waves per tile row
    __shared__ T Xs[2][BK][TS + LDS_PAD];
    __shared__ T Ys[2][BK][TS + LDS_PAD];
    __shared__ int relr[TS];
    __shared__ int relc[TS];

    const Tile tl = tiles[blockIdx.x];
    const UpdDesc d = descs[tl.desc];
    const int i0 = tl.ti * TS, j0 = tl.tj * TS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave % WPR, wc = wave / WPR;

    if (MODE == MODE_SCATTER) {
        if (tid < TS) {
            int i = i0 + tid;
            int r = -1;
            if (i < d.M) {
                int g = Ls[d.ls_off + i];
                if (g < d.base + d.t_nscol) {
                    r = g - d.base;
                } else {
                    int lo = d.t_psi + d.t_nscol, hi = d.t_psi + d.t_nsrow - 1;
                    while (lo < hi) {
                        int mid = (lo + hi) >> 1;
                        if (Ls[mid] < g) lo = mid + 1; else hi = mid;
                    }
                    r = lo - d.t_psi;
                }
            }
            relr[tid] = r;
        } else if (tid < 2 * TS) {
            int j = j0 + tid - TS;
            relc[tid - TS] = (j < d.N) ? (Ls[d.ls_off + j] - d.base) : -1;
        }
    }

    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = acc_t{0, 0, 0, 0};

    const T *__restrict__ Xp = W + d.x_off;
    const T *__restrict__ Yp = W + d.y_off;
    // thread t stages row r = t % TS of chunk columns k = t/TS + KG*q, q < PER
    const int r = tid % TS, kq = tid / TS;
    const bool rowy = (i0 + r) < d.M, rowx = (j0 + r) < d.N;
    const T *yrow = Yp + (i0 + r);
    const T *xrow = Xp + (j0 + r);
    T ry[PER], rx[PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int gk = k0 + kq + KG * q;
            const bool kin = gk < d.K;
            ry[q] = (kin && rowy) ? yrow[(long long)gk * d.ldy] : (T)0;
            rx[q] = (kin && rowx) ? xrow[(long long)gk * d.ldx] : (T)0;
        }
    };
    fetch(0);
    int buf = 0;
    for (int k0 = 0; k0 < d.K; k0 += BK) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            Ys[buf][kq + KG * q][r] = ry[q];
            Xs[buf][kq + KG * q][r] = rx[q];
        }
        __syncthreads();
        if (k0 + BK < d.K) fetch(k0 + BK);      // next chunk in flight during the MFMAs
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T xa[2], yb[2];
#pragma unroll
            for (int m = 0; m < 2; m++) xa[m] = Xs[buf][kk + (lane >> 4)][wc * 32 + m * 16 + (lane & 15)];
#pragma unroll
            for (int n = 0; n < 2; n++) yb[n] = Ys[buf][kk + (lane >> 4)][wr * 32 + n * 16 + (lane & 15)];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++) acc[m][n] = MfmaOp<T>::mma(xa[m], yb[n], acc[m][n]);
        }
        buf ^= 1;
    }
    if (MODE == MODE_STORE) __syncthreads();   // in-place TRSM: every read of the panel is done

    // epilogue: acc[m][n][v] = C(i, j), i = i0 + 32wr + 16n + lane%16,
    //                                   j = j0 + 32wc + 16m + row(lane, v)
#pragma unroll
    for (int m = 0; m < 2; m++) {
#pragma unroll
        for (int n = 0; n < 2; n++) {
            const int il = wr * 32 + n * 16 + (lane & 15);
            const int i = i0 + il;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int jl = wc * 32 + m * 16 + MfmaOp<T>::row(lane, v);
                const int j = j0 + jl;
                if (i >= d.M || j >= d.N) continue;
                const T c = acc[m][n][v];
                if (MODE == MODE_SCATTER) {
                    if (i < j) continue;
                    atomic_sub(W + d.c_off + (long long)relc[jl] * d.ldc + relr[il], c);
                } else if (MODE == MODE_TRAIL) {
                    if (i < j) continue;
                    W[d.c_off + (long long)j * d.ldc + i] -= c;
                } else {
                    W[d.c_off + (long long)j * d.ldc + i] = c;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal block Cholesky + triangular inverse, one 256-thread workgroup per
// 64x64 block held in LDS (a block narrower than 64 is extended by the
// identity).  Right-looking, two barriers per column.  The lower factor is
// written back in place (the strict upper triangle is left alone, as dpotrf
// leaves it) and inv(L) goes to the block's inverse slot (ld nb).  A
// non-positive pivot at column j records col0 + j in *fail (atomic min).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(const DiagTask *__restrict__ tasks,
                                                    T *__restrict__ W, int *__restrict__ fail) {
    __shared__ T A[NB][NB + 1];
    __shared__ T X[NB][NB + 1];
    __shared__ T red[4][NB];
    __shared__ T dsh[NB];
    __shared__ int sfail;
    const DiagTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const int nb = t.nb;
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e & (NB - 1), c = e >> 6;
        T v = 0;
        if (i < nb && c < nb) {
            if (c <= i) v = W[t.off + i + (long long)c * t.ld];
        } else if (i == c) {
            v = 1;
        }
        A[i][c] = v;
    }
    if (tid == 0) sfail = -1;
    const int ri = tid >> 2, cg = tid & 3;
    for (int j = 0; j < NB; j++) {
        __syncthreads();
        T djj = A[j][j];
        if (!(djj > (T)0)) {
            if (tid == 0 && j < nb && sfail < 0) sfail = j;
            djj = 1;
        }
        const T dj = sqrt(djj);
        if (tid > j && tid < NB) A[tid][j] /= dj;
        if (tid == 0) dsh[j] = dj;
        __syncthreads();
        if (ri > j) {
            const T lij = A[ri][j];
            for (int c = j + 1 + cg; c <= ri; c += 4) A[ri][c] -= lij * A[c][j];
        }
    }
    __syncthreads();
    if (tid < NB) A[tid][tid] = dsh[tid];
    if (tid == 0 && sfail >= 0) atomicMin(fail, t.col0 + sfail);
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e & (NB - 1), c = e >> 6;
        if (i < nb && c < nb && c <= i) W[t.off + i + (long long)c * t.ld] = A[i][c];
    }
    // inverse, row by row: X[i][c] = -(sum_{k=c}^{i-1} L[i][k] X[k][c]) / L[i][i]
    const int c = tid & (NB - 1), part = tid >> 6;
    for (int i = 0; i < NB; i++) {
        T s = 0;
        for (int k = c + part; k < i; k += 4) s += A[i][k] * X[k][c];
        red[part][c] = s;
        __syncthreads();
        if (part == 0) {
            const T lii = A[i][i];
            X[i][c] = (c < i) ? -(red[0][c] + red[1][c] + red[2][c] + red[3][c]) / lii
                              : ((c == i) ? (T)1 / lii : (T)0);
        }
        __syncthreads();
    }
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e & (NB - 1), cc = e >> 6;
        if (i < nb && cc < nb) W[t.winv + i + (long long)cc * nb] = X[i][cc];
    }
}

}  // namespace mxc
