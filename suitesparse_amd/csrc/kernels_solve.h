// kernels_solve.h -- device code of the supernodal triangular solves, the
// permutations, the FP64 residual SpMV and the refinement helpers (included
// once, by mxc_device.hip).
//
// Replaces CHOLMOD/Supernodal/t_cholmod_super_solve_worker.c (lsolve :20-303,
// ltsolve :309-580; the reference has no device solve) and the permutation
// steps of CHOLMOD/Cholesky/t_cholmod_psolve_worker.c:40-110.  Diagonal
// blocks are applied through their inverses (kept from the factorization),
// so every step is a small dense mat-vec; off-diagonal rows of a supernode are
// split across workgroups, and one launch carries one 64-column step of every
// supernode of a level.  X is column-major n x nrhs (ld n), permuted order.
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

// Y[i] = (To) X[i] (grid-stride)
template <typename Ti, typename To>
__global__ void k_cast(const Ti *__restrict__ X, long long len, To *__restrict__ Y) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (long long)gridDim.x * blockDim.x)
        Y[i] = (To)X[i];
}

// Cross-lane moves for the wave butterfly, all on the VALU (no LDS traffic):
// v_permlane32_swap / v_permlane16_swap exchange the upper half of one operand
// with the lower half of the other across 32- / 16-lane halves, and DPP row
// operations move data inside a 16-lane row.  64-bit values move as two words.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// a + b after swapping across 32-lane (W = 32) or 16-lane (W = 16) halves: in
// a lane of the lower half the sum of its own a and its partner's a, in a lane
// of the upper half the sum of its partner's b and its own b.
template <int W>
__device__ __forceinline__ void swap_words(unsigned &a, unsigned &b) {
    if constexpr (W == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    }
}
template <int W>
__device__ __forceinline__ float swap_sum(float a, float b) {
    unsigned x = __float_as_uint(a), y = __float_as_uint(b);
    swap_words<W>(x, y);
    return __uint_as_float(x) + __uint_as_float(y);
}
template <int W>
__device__ __forceinline__ double swap_sum(double a, double b) {
    const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
    const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
    unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32);
    unsigned blo = (unsigned)ub, bhi = (unsigned)(ub >> 32);
    swap_words<W>(alo, blo);
    swap_words<W>(ahi, bhi);
    const double x = __longlong_as_double((long long)(((unsigned long long)ahi << 32) | alo));
    const double y = __longlong_as_double((long long)(((unsigned long long)bhi << 32) | blo));
    return x + y;
}

// One halving step: lanes with bit 4*HALF set keep the upper HALF values,
// the others the lower HALF, each summed with its partner's copy (a lane that
// differs in that bit; after the last step the sums over each group of 4
// lanes cover all 64).
template <int HALF, typename T>
__device__ __forceinline__ void fold_half(T (&v)[16], int lane) {
    if constexpr (HALF == 8 || HALF == 4) {
#pragma unroll
        for (int i = 0; i < HALF; i++) v[i] = swap_sum<4 * HALF>(v[i], v[i + HALF]);
    } else {
        constexpr int MASK = 4 * HALF;
        const bool up = (lane & MASK) != 0;
#pragma unroll
        for (int i = 0; i < HALF; i++) {
            const T keep = up ? v[i + HALF] : v[i];
            const T send = up ? v[i] : v[i + HALF];
            // partner across lane bit 3: row_ror:8 (lane ^ 8); across bit 2:
            // row_half_mirror (lane i <-> 7 - i of each 8 lanes).  One uniform DPP
            // move each: a per-lane choice between two moves is compiled into
            // divergent branches, where DPP reads of disabled lanes return 0.
            const T got = (HALF == 2) ? dpp_mov<0x128>(send) : dpp_mov<0x141>(send);
            v[i] = keep + got;
        }
    }
}

// Sum of 16 per-lane values over the 64 lanes of a wave, transposed: on
// return lane l holds the total of index (l >> 2) & 15 (every lane of a
// group of 4 holds the same total).
template <typename T>
__device__ __forceinline__ T wave_sum16(T (&v)[16], int lane) {
    fold_half<8>(v, lane);
    fold_half<4>(v, lane);
    fold_half<2>(v, lane);
    fold_half<1>(v, lane);
    T s = v[0];
    s += dpp_mov<0x4E>(s);      // quad_perm [2,3,0,1]: lane ^ 2
    s += dpp_mov<0xB1>(s);      // quad_perm [1,0,3,2]: lane ^ 1
    return s;
}

// Right-hand sides per pass of a sweep workgroup.
constexpr int ZC = 8;

// One launch = one 64-column step of every supernode of a level.  The
// forward sweep of supernode s (t_cholmod_super_solve_worker.c:20-303: dtrsv
// on the diagonal block, dgemv below it, scatter) is pipelined by one block:
//   diag workgroups, block t:   x_t = inv(L_tt) (x_t - L[t, t-1] x_{t-1})
//   off  workgroups, block t-1: x[g(r)] -= L[r, t-1] x_{t-1}  for the rows r
//                               below block t (r >= nscol: ancestor rows, Ls)
// Both read x_{t-1}, final after the previous launch; diag writes block t's
// rows, off writes the rows below it, so no two workgroups of a launch touch
// one row, except ancestor rows shared by supernodes of one level (atomics).
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_fwd(const SweepTask *__restrict__ tasks, int ndiag,
                                                   const T *__restrict__ W,
                                                   const int *__restrict__ Ls, T *__restrict__ X,
                                                   long long n, int nrhs) {
    __shared__ T xs[ZC][NB];
    __shared__ T red[4][ZC][NB];
    const SweepTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    T *xb = X + t.k1 + t.c0;
    if ((int)blockIdx.x < ndiag) {
        const int nb = t.nb, nbp = t.r0;
        // lane i: row c0 + i of block t; wave w: columns 16w .. 16w+15
        const T *Lb = W + t.lx + (long long)(t.c0 - NB) * t.nsrow + t.c0 + lane;
        T lp[16], iv[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int k = 16 * w + j;
            lp[j] = (lane < nb && k < nbp) ? Lb[(long long)k * t.nsrow] : (T)0;
            iv[j] = (lane < nb && k < nb) ? W[t.winv + lane + (long long)k * nb] : (T)0;
        }
        for (int z0 = 0; z0 < nrhs; z0 += ZC) {
            const int zc = min(ZC, nrhs - z0);
            for (int e = tid; e < ZC * NB; e += 256) {
                const int z = e >> 6, k = e & 63;
                xs[z][k] = (z < zc && k < nbp) ? xb[(long long)(z0 + z) * n + k - NB] : (T)0;
            }
            __syncthreads();
#pragma unroll
            for (int z = 0; z < ZC; z++) {
                T s = 0;
#pragma unroll
                for (int j = 0; j < 16; j++) s += lp[j] * xs[z][16 * w + j];
                red[w][z][lane] = s;
            }
            __syncthreads();
            for (int z = w; z < ZC; z += 4) {
                T v = 0;
                if (z < zc && lane < nb)
                    v = xb[(long long)(z0 + z) * n + lane] -
                        ((red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]));
                xs[z][lane] = v;
            }
            __syncthreads();
#pragma unroll
            for (int z = 0; z < ZC; z++) {
                T s = 0;
#pragma unroll
                for (int j = 0; j < 16; j++) s += iv[j] * xs[z][16 * w + j];
                red[w][z][lane] = s;
            }
            __syncthreads();
            for (int z = w; z < zc; z += 4)
                if (lane < nb)
                    xb[(long long)(z0 + z) * n + lane] =
                        (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]);
            __syncthreads();
        }
        return;
    }
    // off: one thread per row, the row's nb values loaded up front
    const int nb = t.nb;
    const int r = t.r0 + tid;
    const bool active = r < t.r1;
    long long g = 0;
    if (active) g = (r < t.nscol) ? (long long)(t.k1 + r) : (long long)Ls[t.psi + r];
    const T *Lrow = W + t.lx + (long long)t.c0 * t.nsrow + r;
    T l[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) l[k] = (active && k < nb) ? Lrow[(long long)k * t.nsrow] : (T)0;
    for (int z0 = 0; z0 < nrhs; z0 += ZC) {
        const int zc = min(ZC, nrhs - z0);
        for (int e = tid; e < ZC * NB; e += 256) {
            const int z = e >> 6, k = e & 63;
            xs[z][k] = (z < zc && k < nb) ? xb[(long long)(z0 + z) * n + k] : (T)0;
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int z = 0; z < ZC; z++) {
                if (z < zc) {
                    T s0 = 0, s1 = 0;
#pragma unroll
                    for (int k = 0; k < NB; k += 2) {
                        s0 += l[k] * xs[z][k];
                        s1 += l[k + 1] * xs[z][k + 1];
                    }
                    T *dst = X + g + (long long)(z0 + z) * n;
                    if (r < t.nscol) *dst -= s0 + s1;
                    else atomic_add(dst, -(s0 + s1));
                }
            }
        }
        __syncthreads();
    }
}

// Columns [c, c + 16) of one row of a supernode (zero beyond nb).
template <typename T>
__device__ __forceinline__ void load_row16(T (&lc)[16], const T *Lrow, int c, int nb, int nsrow,
                                           bool active) {
#pragma unroll
    for (int j = 0; j < 16; j++)
        lc[j] = (active && c + j < nb) ? Lrow[(long long)(c + j) * nsrow] : (T)0;
}

// dst[q] = sum over the wave's lanes of lc[q] * x  (q < 16, written by lanes 4q)
template <typename T>
__device__ __forceinline__ void col_sums16(const T (&lc)[16], T x, T *dst, int lane) {
    T v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) v[j] = lc[j] * x;
    const T s = wave_sum16(v, lane);
    if ((lane & 3) == 0) dst[(lane >> 2) & 15] = s;
}

// Backward sweep (t_cholmod_super_solve_worker.c:309-580: gather, dgemv with
// L', dtrsv with L_tt'), pipelined the same way, steps in descending order:
//   diag workgroups, block t:   x_t = inv(L_tt)' (x_t - sum of partials_t
//                                                 - L[t+1, t]' x_{t+1})
//   off  workgroups, block t-1: partial = L[r, t-1]' x[g(r)] over their rows
//                               r below block t (one slot per workgroup)
// Partials are written, not accumulated: no atomics, and the order of the
// sums is fixed, so the sweep is deterministic.
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_bwd(const SweepTask *__restrict__ tasks, int ndiag,
                                                   const T *__restrict__ W,
                                                   const int *__restrict__ Ls, T *__restrict__ X,
                                                   T *__restrict__ part, long long n, int nrhs) {
    __shared__ T xs[ZC][NB];
    __shared__ T red[4][ZC][NB];
    const SweepTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int qcol = (lane >> 2) & 15;          // column of a wave_sum16 total
    T *xb = X + t.k1 + t.c0;
    if ((int)blockIdx.x < ndiag) {
        // The two products with transposed operands (L[t+1, t]' and inv(L_tt)')
        // go through a transposed LDS copy: lane c then owns output column c and
        // sums its wave's 16 rows with plain LDS reads (no lane exchanges).
        __shared__ T tt[NB][NB + 1];         // tt[c][k]: transposed 64 x 64 block
        __shared__ T ptot[ZC][NB];           // partial sums of the rows below
        const int nb = t.nb, nbn = t.r0;
        // lane k: row c0 + 64 + k (block t+1) / row k of the inverse;
        // wave w: columns 16w .. 16w+15 of block t
        const T *Lb = W + t.lx + (long long)t.c0 * t.nsrow + t.c0 + NB + lane;
        T ln[16], iv[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int k = 16 * w + j;
            ln[j] = (lane < nbn && k < nb) ? Lb[(long long)k * t.nsrow] : (T)0;
            iv[j] = (lane < nb && k < nb) ? W[t.winv + lane + (long long)k * nb] : (T)0;
        }
        const long long pstride = (long long)nrhs * NB;
        for (int z0 = 0; z0 < nrhs; z0 += ZC) {
            const int zc = min(ZC, nrhs - z0);
            for (int e = tid; e < ZC * NB; e += 256) {
                const int z = e >> 6, k = e & 63;
                xs[z][k] = (z < zc && k < nbn) ? xb[(long long)(z0 + z) * n + NB + k] : (T)0;
            }
#pragma unroll
            for (int j = 0; j < 16; j++) tt[16 * w + j][lane] = ln[j];
            // partial sums of the rows below: wave w takes slots p = w (mod 4), every
            // right-hand side of the chunk, two slots in flight
            {
                T s0[ZC], s1[ZC];
#pragma unroll
                for (int z = 0; z < ZC; z++) s0[z] = s1[z] = 0;
                const T *pp = part + ((long long)t.p0 * nrhs + z0) * NB + lane;
                int p = w;
                for (; p + 4 < t.np; p += 8) {
#pragma unroll
                    for (int z = 0; z < ZC; z++) {
                        if (z < zc) {
                            s0[z] += pp[(long long)p * pstride + z * NB];
                            s1[z] += pp[(long long)(p + 4) * pstride + z * NB];
                        }
                    }
                }
                if (p < t.np) {
#pragma unroll
                    for (int z = 0; z < ZC; z++)
                        if (z < zc) s0[z] += pp[(long long)p * pstride + z * NB];
                }
#pragma unroll
                for (int z = 0; z < ZC; z++) red[w][z][lane] = s0[z] + s1[z];
            }
            __syncthreads();
            // ptot = partials; red = neighbour term L[t+1, t]' x_{t+1}, wave w: rows 16w..
            for (int z = w; z < ZC; z += 4)
                ptot[z][lane] = (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]);
            __syncthreads();
#pragma unroll
            for (int z = 0; z < ZC; z++) {
                T a = 0;
                if (z < zc) {
#pragma unroll
                    for (int j = 0; j < 16; j++) a += tt[lane][16 * w + j] * xs[z][16 * w + j];
                }
                red[w][z][lane] = a;
            }
            __syncthreads();
            // v = x_t - partials - neighbour term (into xs); the inverse goes into tt
            for (int z = w; z < ZC; z += 4) {
                T v = 0;
                if (z < zc && lane < nb)
                    v = xb[(long long)(z0 + z) * n + lane] - ptot[z][lane] -
                        ((red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]));
                xs[z][lane] = v;
            }
#pragma unroll
            for (int j = 0; j < 16; j++) tt[16 * w + j][lane] = iv[j];
            __syncthreads();
            // x_t = inv(L_tt)' v: column c sums inv[k][c] v[k] over wave w's rows k
#pragma unroll
            for (int z = 0; z < ZC; z++) {
                T a = 0;
                if (z < zc) {
#pragma unroll
                    for (int j = 0; j < 16; j++) a += tt[lane][16 * w + j] * xs[z][16 * w + j];
                }
                red[w][z][lane] = a;
            }
            __syncthreads();
            for (int z = w; z < zc; z += 4)
                if (lane < nb)
                    xb[(long long)(z0 + z) * n + lane] =
                        (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]);
            __syncthreads();
        }
        return;
    }
    // off: one thread per row, transposed wave sums, one partial per workgroup.
    // The row is held as four 16-column chunks so that every register index
    // is a compile-time constant.
    const int nb = t.nb;
    const int r = t.r0 + tid;
    const bool active = r < t.r1;
    long long g = 0;
    if (active) g = (r < t.nscol) ? (long long)(t.k1 + r) : (long long)Ls[t.psi + r];
    const T *Lrow = W + t.lx + (long long)t.c0 * t.nsrow + r;
    T l0[16], l1[16], l2[16], l3[16];
    load_row16(l0, Lrow, 0, nb, t.nsrow, active);
    load_row16(l1, Lrow, 16, nb, t.nsrow, active);
    load_row16(l2, Lrow, 32, nb, t.nsrow, active);
    load_row16(l3, Lrow, 48, nb, t.nsrow, active);
    for (int z0 = 0; z0 < nrhs; z0 += ZC) {
        const int zc = min(ZC, nrhs - z0);
        T xr[ZC];
#pragma unroll
        for (int z = 0; z < ZC; z++) xr[z] = (active && z < zc) ? X[g + (long long)(z0 + z) * n] : (T)0;
#pragma unroll
        for (int z = 0; z < ZC; z++) {
            if (z < zc) {
                col_sums16(l0, xr[z], &red[w][z][0], lane);
                col_sums16(l1, xr[z], &red[w][z][16], lane);
                col_sums16(l2, xr[z], &red[w][z][32], lane);
                col_sums16(l3, xr[z], &red[w][z][48], lane);
            }
        }
        __syncthreads();
        for (int z = w; z < zc; z += 4)
            part[((long long)t.p0 * nrhs + z0 + z) * NB + lane] =
                (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Small supernodes (nscol <= 64 and at most SMALL_ROWS rows below): one wave
// solves one supernode completely (diagonal block through its inverse, then
// the rows below), four supernodes per workgroup, so the wide bottom levels
// of a forest need one workgroup per four supernodes and one launch per level.
// Waves share no data; the kernels have no workgroup barrier.
// ---------------------------------------------------------------------------
constexpr int SMALL_ROWS = 128;
constexpr int ZS = 4;          // right-hand sides per pass

template <typename T>
__global__ __launch_bounds__(256) void k_sweep_small_fwd(const SweepTask *__restrict__ tasks, int ntask,
                                                         const T *__restrict__ W,
                                                         const int *__restrict__ Ls, T *__restrict__ X,
                                                         long long n, int nrhs) {
    __shared__ T xs[4][ZS][NB];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ti = blockIdx.x * 4 + w;
    if (ti >= ntask) return;
    const SweepTask t = tasks[ti];
    const int nb = t.nb;
    T *xb = X + t.k1;
    for (int z0 = 0; z0 < nrhs; z0 += ZS) {
        const int zc = min(ZS, nrhs - z0);
#pragma unroll
        for (int z = 0; z < ZS; z++)
            xs[w][z][lane] = (z < zc && lane < nb) ? xb[(long long)(z0 + z) * n + lane] : (T)0;
        wave_sync_lds();
        // y = inv(L_bb) x_b: lane i forms row i
        T y[ZS] = {};
        for (int kc = 0; kc < nb; kc += 16) {
            T iv[16];
#pragma unroll
            for (int j = 0; j < 16; j++)
                iv[j] = (lane < nb && kc + j <= lane) ? W[t.winv + lane + (long long)(kc + j) * nb] : (T)0;
#pragma unroll
            for (int z = 0; z < ZS; z++)
#pragma unroll
                for (int j = 0; j < 16; j++) y[z] += iv[j] * xs[w][z][(kc + j) & 63];
        }
        wave_sync_lds();
#pragma unroll
        for (int z = 0; z < ZS; z++) {
            if (z < zc && lane < nb) xb[(long long)(z0 + z) * n + lane] = y[z];
            xs[w][z][lane] = y[z];
        }
        wave_sync_lds();
        // rows below (ancestor rows, shared with other supernodes: atomics)
        for (int r0 = nb; r0 < t.nsrow; r0 += 64) {
            const int r = r0 + lane;
            const bool active = r < t.nsrow;
            const T *Lrow = W + t.lx + r;
            T acc[ZS] = {};
            for (int kc = 0; kc < nb; kc += 16) {
                T l[16];
#pragma unroll
                for (int j = 0; j < 16; j++)
                    l[j] = (active && kc + j < nb) ? Lrow[(long long)(kc + j) * t.nsrow] : (T)0;
#pragma unroll
                for (int z = 0; z < ZS; z++)
#pragma unroll
                    for (int j = 0; j < 16; j++) acc[z] += l[j] * xs[w][z][(kc + j) & 63];
            }
            if (active) {
                const long long g = Ls[t.psi + r];
#pragma unroll
                for (int z = 0; z < ZS; z++)
                    if (z < zc) atomic_add(X + g + (long long)(z0 + z) * n, -acc[z]);
            }
        }
        wave_sync_lds();
    }
}

// Tiny supernodes (at most TINY_COLS columns and TINY_ROWS rows): one lane
// solves a whole supernode, so the wide bottom levels of a forest (mostly
// one-column supernodes with a few rows below) take one lane each instead of
// a wave each.  A lane walks its rows one by one, so the row limit is low: a
// supernode with more rows keeps a wave, which takes 64 rows at once.  Same
// data as the small kernels: the inverse of the diagonal block at winv
// (element (i, c) at i + c nb), L[r, c] at lx + r + c nsrow.
constexpr int TINY_COLS = 8;
constexpr int TINY_ROWS = 8;

// forward: x_b = inv(L_bb) x_b, then x[g(r)] -= L[r, :] x_b for the rows below
// (ancestor rows, shared with other supernodes: atomics)
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_tiny_fwd(const SweepTask *__restrict__ tasks, int ntask,
                                                        const T *__restrict__ W,
                                                        const int *__restrict__ Ls, T *__restrict__ X,
                                                        long long n, int nrhs) {
    const int ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= ntask) return;
    const SweepTask t = tasks[ti];
    const int nb = t.nb;
    const T *inv = W + t.winv;
    const T *Lb = W + t.lx;
    for (int z = 0; z < nrhs; z++) {
        T *xz = X + (long long)z * n;
        T xb[TINY_COLS], y[TINY_COLS];
#pragma unroll
        for (int c = 0; c < TINY_COLS; c++) xb[c] = c < nb ? xz[t.k1 + c] : (T)0;
#pragma unroll
        for (int i = 0; i < TINY_COLS; i++) {
            T a = 0;
#pragma unroll
            for (int c = 0; c <= i; c++)
                if (i < nb) a += inv[i + c * nb] * xb[c];
            y[i] = a;
            if (i < nb) xz[t.k1 + i] = a;
        }
        for (int r = nb; r < t.nsrow; r++) {
            T a = 0;
#pragma unroll
            for (int c = 0; c < TINY_COLS; c++)
                if (c < nb) a += Lb[r + (long long)c * t.nsrow] * y[c];
            atomic_add(xz + Ls[t.psi + r], -a);
        }
    }
}

// backward: x_b = inv(L_bb)' (x_b - L[below, :]' x[g(below)]); the rows below
// are ancestor rows, final before the launch
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_tiny_bwd(const SweepTask *__restrict__ tasks, int ntask,
                                                        const T *__restrict__ W,
                                                        const int *__restrict__ Ls, T *__restrict__ X,
                                                        long long n, int nrhs) {
    const int ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= ntask) return;
    const SweepTask t = tasks[ti];
    const int nb = t.nb;
    const T *inv = W + t.winv;
    const T *Lb = W + t.lx;
    for (int z = 0; z < nrhs; z++) {
        T *xz = X + (long long)z * n;
        T v[TINY_COLS];
#pragma unroll
        for (int c = 0; c < TINY_COLS; c++) v[c] = c < nb ? xz[t.k1 + c] : (T)0;
        for (int r = nb; r < t.nsrow; r++) {
            const T xr = xz[Ls[t.psi + r]];
#pragma unroll
            for (int c = 0; c < TINY_COLS; c++)
                if (c < nb) v[c] -= Lb[r + (long long)c * t.nsrow] * xr;
        }
#pragma unroll
        for (int c = 0; c < TINY_COLS; c++) {
            T a = 0;
#pragma unroll
            for (int i = c; i < TINY_COLS; i++)
                if (i < nb) a += inv[i + c * nb] * v[i];
            if (c < nb) xz[t.k1 + c] = a;
        }
    }
}

// Quarter-wave supernodes (more than tiny, at most QUAD_COLS columns and
// QUAD_ROWS rows below): sixteen lanes solve one supernode, four per wave and
// sixteen per workgroup, where a wave of its own would leave 48 lanes idle.
// Lane q owns row q of the diagonal block and row nb + q below it; the
// inverse's row (forward) or column (backward) and the L entries of its row
// below are loaded once for every right-hand side.  Groups share no data
// (each has its own LDS slice); synchronization is per wave.
constexpr int QUAD_COLS = 16;
constexpr int QUAD_ROWS = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_sweep_quad_fwd(const SweepTask *__restrict__ tasks, int ntask,
                                                        const T *__restrict__ W,
                                                        const int *__restrict__ Ls, T *__restrict__ X,
                                                        long long n, int nrhs) {
    __shared__ T ys[16][QUAD_COLS];
    const int q = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int ti = blockIdx.x * 16 + grp;
    const bool on = ti < ntask;
    SweepTask t = {};
    if (on) t = tasks[ti];
    const int nb = on ? t.nb : 0, below = on ? t.nsrow - t.nb : 0;
    T iv[QUAD_COLS], lr[QUAD_COLS];
#pragma unroll
    for (int c = 0; c < QUAD_COLS; c++) {
        iv[c] = (q < nb && c <= q) ? W[t.winv + q + (long long)c * nb] : (T)0;
        lr[c] = (q < below && c < nb) ? W[t.lx + nb + q + (long long)c * t.nsrow] : (T)0;
    }
    const long long g = q < below ? (long long)Ls[t.psi + nb + q] : 0;
    for (int z0 = 0; z0 < nrhs; z0 += ZS) {
        // the pass's right-hand-side entries, all loads in flight before the
        // first LDS fence (each fence would otherwise wait for the next load)
        T xv[ZS];
#pragma unroll
        for (int zz = 0; zz < ZS; zz++)
            xv[zz] = (q < nb && z0 + zz < nrhs) ? X[(long long)(z0 + zz) * n + t.k1 + q] : (T)0;
#pragma unroll
        for (int zz = 0; zz < ZS; zz++) {
            if (z0 + zz >= nrhs) break;
            T *xz = X + (long long)(z0 + zz) * n;
            ys[grp][q] = xv[zz];
            wave_sync_lds();
            T y = 0;
#pragma unroll
            for (int c = 0; c < QUAD_COLS; c++) y += iv[c] * ys[grp][c];
            wave_sync_lds();
            if (q < nb) xz[t.k1 + q] = y;
            ys[grp][q] = y;
            wave_sync_lds();
            if (q < below) {
                T a = 0;
#pragma unroll
                for (int c = 0; c < QUAD_COLS; c++) a += lr[c] * ys[grp][c];
                atomic_add(xz + g, -a);
            }
            wave_sync_lds();
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sweep_quad_bwd(const SweepTask *__restrict__ tasks, int ntask,
                                                        const T *__restrict__ W,
                                                        const int *__restrict__ Ls, T *__restrict__ X,
                                                        long long n, int nrhs) {
    __shared__ T vs[16][QUAD_COLS];
    const int q = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int ti = blockIdx.x * 16 + grp;
    const bool on = ti < ntask;
    SweepTask t = {};
    if (on) t = tasks[ti];
    const int nb = on ? t.nb : 0, below = on ? t.nsrow - t.nb : 0;
    // lane q as column q: the inverse's column q and L's column q below the block
    T iv[QUAD_COLS], lc[QUAD_ROWS];
#pragma unroll
    for (int i = 0; i < QUAD_COLS; i++) iv[i] = (q < nb && i >= q && i < nb) ? W[t.winv + i + (long long)q * nb] : (T)0;
#pragma unroll
    for (int r = 0; r < QUAD_ROWS; r++) lc[r] = (q < nb && r < below) ? W[t.lx + nb + r + (long long)q * t.nsrow] : (T)0;
    const long long g = q < below ? (long long)Ls[t.psi + nb + q] : 0;
    for (int z0 = 0; z0 < nrhs; z0 += ZS) {
        // the pass's entries (ancestor rows: final before the launch), all
        // loads in flight before the first LDS fence
        T xa[ZS], xo[ZS];
#pragma unroll
        for (int zz = 0; zz < ZS; zz++) {
            const bool zin = z0 + zz < nrhs;
            xa[zz] = (q < below && zin) ? X[(long long)(z0 + zz) * n + g] : (T)0;
            xo[zz] = (q < nb && zin) ? X[(long long)(z0 + zz) * n + t.k1 + q] : (T)0;
        }
#pragma unroll
        for (int zz = 0; zz < ZS; zz++) {
            if (z0 + zz >= nrhs) break;
            T *xz = X + (long long)(z0 + zz) * n;
            vs[grp][q] = xa[zz];
            wave_sync_lds();
            T p = 0;
#pragma unroll
            for (int r = 0; r < QUAD_ROWS; r++) p += lc[r] * vs[grp][r];
            const T v = q < nb ? xo[zz] - p : (T)0;
            wave_sync_lds();
            vs[grp][q] = v;
            wave_sync_lds();
            T y = 0;
#pragma unroll
            for (int i = 0; i < QUAD_COLS; i++) y += iv[i] * vs[grp][i];
            if (q < nb) xz[t.k1 + q] = y;
            wave_sync_lds();
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sweep_small_bwd(const SweepTask *__restrict__ tasks, int ntask,
                                                         const T *__restrict__ W,
                                                         const int *__restrict__ Ls, T *__restrict__ X,
                                                         long long n, int nrhs) {
    __shared__ T ps[4][ZS][NB];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ti = blockIdx.x * 4 + w;
    if (ti >= ntask) return;
    const SweepTask t = tasks[ti];
    const int nb = t.nb;
    const int qcol = (lane >> 2) & 15;
    T *xb = X + t.k1;
    for (int z0 = 0; z0 < nrhs; z0 += ZS) {
        const int zc = min(ZS, nrhs - z0);
        // p_k = sum over the rows below of L[r, k] x[g(r)], 16 columns and two
        // right-hand sides at a time (fewer live accumulators: higher occupancy;
        // the second pair re-reads the same L rows from cache)
        for (int kc = 0; kc < nb; kc += 16)
            for (int zp = 0; zp < zc; zp += 2) {
                T v[2][16] = {};
                for (int r0 = nb; r0 < t.nsrow; r0 += 64) {
                    const int r = r0 + lane;
                    const bool active = r < t.nsrow;
                    const T *Lrow = W + t.lx + r;
                    T xr[2];
                    const long long g = active ? (long long)Ls[t.psi + r] : 0;
#pragma unroll
                    for (int z = 0; z < 2; z++)
                        xr[z] = (active && zp + z < zc) ? X[g + (long long)(z0 + zp + z) * n] : (T)0;
#pragma unroll
                    for (int j = 0; j < 16; j++) {
                        const T l = (active && kc + j < nb) ? Lrow[(long long)(kc + j) * t.nsrow] : (T)0;
#pragma unroll
                        for (int z = 0; z < 2; z++) v[z][j] += l * xr[z];
                    }
                }
#pragma unroll
                for (int z = 0; z < 2; z++) {
                    const T s = wave_sum16(v[z], lane);
                    if ((lane & 3) == 0) ps[w][zp + z][(kc + qcol) & 63] = s;
                }
            }
        wave_sync_lds();
        // v = x_b - p, then x_b = inv(L_bb)' v: lane k forms entry k
#pragma unroll
        for (int z = 0; z < ZS; z++) {
            const T xv = (z < zc && lane < nb) ? xb[(long long)(z0 + z) * n + lane] : (T)0;
            const T pv = (lane < nb && nb < t.nsrow) ? ps[w][z][lane] : (T)0;
            ps[w][z][lane] = xv - pv;
        }
        wave_sync_lds();
        T y[ZS] = {};
        for (int ic = 0; ic < nb; ic += 16) {
            T iv[16];
#pragma unroll
            for (int j = 0; j < 16; j++)
                iv[j] = (lane < nb && ic + j < nb && ic + j >= lane) ? W[t.winv + (ic + j) + (long long)lane * nb]
                                                                     : (T)0;
#pragma unroll
            for (int z = 0; z < ZS; z++)
#pragma unroll
                for (int j = 0; j < 16; j++) y[z] += iv[j] * ps[w][z][(ic + j) & 63];
        }
#pragma unroll
        for (int z = 0; z < ZS; z++)
            if (z < zc && lane < nb) xb[(long long)(z0 + z) * n + lane] = y[z];
        wave_sync_lds();
    }
}

// ---------------------------------------------------------------------------
// Chained sweeps: one launch per level for the supernodes that are not small.
// Workgroup (s, t) computes block t of supernode s.  It accumulates the
// products with the blocks it depends on as they become final and applies the
// inverse of the diagonal block at the end.  A per-supernode counter of
// finished blocks passes blocks between the workgroups of a launch: block
// values and counters are read and written with device-scope atomic loads and
// stores (performed at the device coherence point, no cache maintenance), and
// a producer waits for its block's stores to complete (s_waitcnt) before the
// workgroup barrier that precedes the counter store; a consumer issues its
// block loads only after its counter load returned.  The launch order puts
// every block after the blocks it depends on, so a workgroup waits only for
// workgroups of lower index: the lowest unfinished workgroup never waits, and
// workgroups start in index order, so every wait ends.  A wait is still
// bounded (CHAIN_SPIN polls); past that it raises *abort, every wait of the
// launch returns, and the host reports a device error instead of hanging.
//
// Inside a workgroup, lane l owns output row (forward) or column (backward)
// l of the 64-wide block and wave w the slice 16w .. 16w+15 of the summed
// dimension; the 16 x ZC values of a slice are loaded once per wave (one per
// lane) and broadcast with readlane, and the four waves' partial sums meet in
// LDS once per pass.
// ---------------------------------------------------------------------------
constexpr int CHAIN_SPIN = 1 << 22;
constexpr int CHAIN_RHS = 4;   // columns per launch (the host splits wider right-hand sides)

template <typename T>
__device__ __forceinline__ T ld_dev(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_dev(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-uniform wait until *cnt >= need.  Returns the counter value seen.
__device__ __forceinline__ int chain_wait(const int *cnt, int need, int *abort) {
    int c = __builtin_amdgcn_readfirstlane(ld_dev(cnt));
    for (int it = 1; c < need; it++) {
        __builtin_amdgcn_s_sleep(1);
        if ((it & 63) == 0) {
            const int ab = __builtin_amdgcn_readfirstlane(ld_dev(abort));
            if (ab != 0 || it >= CHAIN_SPIN) {
                if (ab == 0) st_dev(abort, 1);
                return need;
            }
        }
        c = __builtin_amdgcn_readfirstlane(ld_dev(cnt));
    }
    asm volatile("" ::: "memory");   // block loads stay after the counter load
    return c;
}

// Publish block count v of a supernode once every thread's block stores are done.
__device__ __forceinline__ void chain_publish(int *cnt, int v) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) st_dev(cnt, v);
}

__global__ void k_chain_reset(int *__restrict__ cnt, int m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) cnt[i] = 0;
}

template <typename T>
__device__ __forceinline__ T lane_get(T v, int l) {
    if constexpr (sizeof(T) == 8) {
        const long long u = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)u, l);
        const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    }
}

// acc[z] += sum_j l[j] q(j + 16 z), q(m) = lane m's value of q (wave-uniform)
template <typename T, int ZC>
__device__ __forceinline__ void fma_bcast(T (&acc)[ZC], const T (&l)[16], T q) {
#pragma unroll
    for (int z = 0; z < ZC; z++)
#pragma unroll
        for (int j = 0; j < 16; j++) acc[z] += l[j] * lane_get(q, j + 16 * z);
}

// Forward.  Task fields: c0 / nb = block t (diag workgroup, pad[0] = 0) or
// rows [r0, r1) below the supernode's columns (pad[0] = 1); p0 = counter of
// the supernode, np = its number of blocks.
//   diag:  x_t = inv(L_tt) (x_t - sum_{b < t} L[t, b] x_b), then publish t + 1
//   below: x[g(r)] -= sum_b L[r, b] x_b  (ancestor rows, atomics)
template <typename T, int ZC>
__global__ __launch_bounds__(256) void k_chain_fwd(const SweepTask *__restrict__ tasks,
                                                   const T *__restrict__ W,
                                                   const int *__restrict__ Ls, T *__restrict__ X,
                                                   long long n, int nrhs, int *__restrict__ cnt,
                                                   int *__restrict__ abort) {
    __shared__ T red[4][ZC][NB];
    __shared__ T vs[ZC][NB];
    const SweepTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const bool below = t.pad[0] != 0;
    const int row = (below ? t.r0 : t.c0) + lane;  // the lane's row of the supernode
    const bool rok = row < (below ? t.r1 : t.c0 + t.nb);
    const int nblk = below ? t.np : t.c0 / NB;     // blocks summed
    int *mycnt = cnt + t.p0;
    const T *Lr = W + t.lx + row + (long long)(16 * w) * t.nsrow;
    // the inverse (diag): row = lane, columns 16w + j; loaded before any wait
    T iv[16];
#pragma unroll
    for (int j = 0; j < 16; j++)
        iv[j] = (!below && lane < t.nb && 16 * w + j < t.nb)
                    ? W[t.winv + lane + (long long)(16 * w + j) * t.nb] : (T)0;
    auto load_blk = [&](T (&l)[16], int b) {
        const int c = b * NB + 16 * w;             // first column of the wave's slice
#pragma unroll
        for (int j = 0; j < 16; j++)
            l[j] = (rok && c + j < t.nscol) ? Lr[(long long)(b * NB + j) * t.nsrow] : (T)0;
    };
    const int qz = lane >> 4, qj = lane & 15;      // the value lane q loads: column qj, rhs qz
    int seen = 0;
    for (int z0 = 0; z0 < nrhs; z0 += ZC) {
        const int zc = min(ZC, nrhs - z0);
        T acc[ZC];
#pragma unroll
        for (int z = 0; z < ZC; z++) acc[z] = 0;
        T lc[16], ln[16];
        if (nblk > 0) load_blk(lc, 0);
        for (int b = 0; b < nblk; b++) {
            if (b + 1 < nblk) load_blk(ln, b + 1);
            if (seen <= b) seen = chain_wait(mycnt, b + 1, abort);
            const int c = b * NB + 16 * w + qj;
            const T q = (qz < zc && c < t.nscol) ? ld_dev(X + t.k1 + c + (long long)(z0 + qz) * n) : (T)0;
            fma_bcast<T, ZC>(acc, lc, q);
#pragma unroll
            for (int j = 0; j < 16; j++) lc[j] = ln[j];
        }
#pragma unroll
        for (int z = 0; z < ZC; z++) red[w][z][lane] = acc[z];
        __syncthreads();
        if (below) {
            for (int z = w; z < zc; z += 4)
                if (rok) {
                    const T sum = (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]);
                    atomic_add(X + Ls[t.psi + row] + (long long)(z0 + z) * n, -sum);
                }
            __syncthreads();                        // red is rewritten by the next pass
            continue;
        }
        for (int z = w; z < ZC; z += 4) {
            T v = 0;
            if (rok && z < zc)
                v = X[t.k1 + row + (long long)(z0 + z) * n] -
                    ((red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]));
            vs[z][lane] = v;
        }
        __syncthreads();
        // x_t = inv(L_tt) v: row = lane, wave w sums the columns 16w .. 16w+15
#pragma unroll
        for (int z = 0; z < ZC; z++) {
            T a = 0;
#pragma unroll
            for (int j = 0; j < 16; j++) a += iv[j] * vs[z][16 * w + j];
            red[w][z][lane] = a;
        }
        __syncthreads();
        for (int z = w; z < zc; z += 4)
            if (rok)
                st_dev(X + t.k1 + row + (long long)(z0 + z) * n,
                       (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]));
        __syncthreads();
    }
    if (!below) chain_publish(mycnt, t.c0 / NB + 1);
}

// Backward.  Below workgroups (pad[0] = 1) come first in a launch: rows
// [r0, r1) below the supernode's columns times block t = c0 / NB,
//   part[np] = L[r0:r1, t]' x[g(r0:r1)]   (one slot; ancestor rows, final)
// then counter p0 += 1.  Diag workgroups (pad[0] = 0): p0 = counter of the
// supernode's blocks, np = its number of blocks, slots [q0, q0 + nq) of block
// t complete once counter bc reaches bt:
//   x_t = inv(L_tt)' (x_t - sum of slots - sum_{b > t} L[b, t]' x_b),
// then publish np - t (blocks finished, counted from the last).  Slots are
// written, not accumulated, and summed in a fixed order: the sweep is
// deterministic.
template <typename T, int ZC>
__global__ __launch_bounds__(256) void k_chain_bwd(const SweepTask *__restrict__ tasks,
                                                   const T *__restrict__ W,
                                                   const int *__restrict__ Ls, T *__restrict__ X,
                                                   T *__restrict__ part, long long n, int nrhs,
                                                   int *__restrict__ cnt, int *__restrict__ abort) {
    __shared__ T red[4][ZC][NB];
    __shared__ T vs[ZC][NB];
    __shared__ T tile[4][16][NB + 1];              // per wave: 16 rows x 64 columns of L
    const SweepTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const bool below = t.pad[0] != 0;
    const int nblk = t.np, tb = t.c0 / NB;
    const int col = t.c0 + lane;                   // the lane's column of the supernode
    const bool cok = lane < t.nb;
    // the inverse (diag): column = lane, rows 16w + j (inv[k][c] at winv + k + c nb)
    T iv[16];
#pragma unroll
    for (int j = 0; j < 16; j++)
        iv[j] = (!below && cok && 16 * w + j < t.nb) ? W[t.winv + 16 * w + j + (long long)lane * t.nb]
                                                     : (T)0;
    // L rows r + 16w .. r + 16w + 15 of the block's 64 columns (zero at or past
    // rend), read in column segments: lane = (rows 4fq .. 4fq + 3, column
    // fc + 16k), so one load covers 16 columns, 32 bytes each.  put() stages
    // them in the wave's tile and get() reads them back with lane = column:
    // l[j] = L[r + 16w + j, c0 + lane].
    const int fq = lane & 3, fc = lane >> 2;
    const T *Lw = W + t.lx + (long long)(t.c0 + fc) * t.nsrow + 16 * w + 4 * fq;
    const long long cs = 16LL * t.nsrow;           // 16 columns
    auto fetch = [&](T (&g)[16], int r, int rend) {
        const int r4 = r + 16 * w + 4 * fq;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool ok = fc + 16 * k < t.nb;
#pragma unroll
            for (int e = 0; e < 4; e++) g[4 * k + e] = (ok && r4 + e < rend) ? Lw[k * cs + r + e] : (T)0;
        }
    };
    auto put = [&](const T (&g)[16]) {
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int e = 0; e < 4; e++) tile[w][4 * fq + e][fc + 16 * k] = g[4 * k + e];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    auto get = [&](T (&l)[16]) {
#pragma unroll
        for (int j = 0; j < 16; j++) l[j] = tile[w][j][lane];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    const int qz = lane >> 4, qj = lane & 15;      // the value lane q loads: row qj, rhs qz
    int seen = 0;
    bool slots_ready = false;
    for (int z0 = 0; z0 < nrhs; z0 += ZC) {
        const int zc = min(ZC, nrhs - z0);
        T acc[ZC];
#pragma unroll
        for (int z = 0; z < ZC; z++) acc[z] = 0;
        T lc[16], gn[16];
        if (below) {
            // rows [r0, r1): ancestor rows, final before the launch
            fetch(gn, t.r0, t.r1);
            put(gn);
            for (int r = t.r0; r < t.r1; r += NB) {
                get(lc);
                if (r + NB < t.r1) fetch(gn, r + NB, t.r1);
                const int rq = r + 16 * w + qj;
                const T q = (qz < zc && rq < t.r1) ? X[Ls[t.psi + rq] + (long long)(z0 + qz) * n] : (T)0;
                fma_bcast<T, ZC>(acc, lc, q);
                if (r + NB < t.r1) put(gn);
            }
#pragma unroll
            for (int z = 0; z < ZC; z++) red[w][z][lane] = acc[z];
            __syncthreads();
            for (int z = w; z < zc; z += 4)
                st_dev(part + ((long long)t.np * nrhs + z0 + z) * NB + lane,
                       (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]));
            __syncthreads();
            continue;
        }
        // the slots of the rows below block t: wave w takes slots w, w + 4, ...
        if (t.nq > 0) {
            if (!slots_ready) {
                chain_wait(cnt + t.bc, t.bt, abort);
                slots_ready = true;
            }
            for (int k = w; k < t.nq; k += 4)
#pragma unroll
                for (int z = 0; z < ZC; z++)
                    if (z < zc) acc[z] += ld_dev(part + ((long long)(t.q0 + k) * nrhs + z0 + z) * NB + lane);
        }
        // blocks nblk-1 .. tb+1 of the supernode's own rows, as they are finished
        if (tb + 1 < nblk) {
            fetch(gn, (nblk - 1) * NB, t.nscol);
            put(gn);
        }
        for (int b = nblk - 1; b > tb; b--) {
            get(lc);
            if (b - 1 > tb) fetch(gn, (b - 1) * NB, t.nscol);
            if (seen < nblk - b) seen = chain_wait(cnt + t.p0, nblk - b, abort);
            const int rq = b * NB + 16 * w + qj;
            const T q = (qz < zc && rq < t.nscol) ? ld_dev(X + t.k1 + rq + (long long)(z0 + qz) * n) : (T)0;
            fma_bcast<T, ZC>(acc, lc, q);
            if (b - 1 > tb) put(gn);
        }
#pragma unroll
        for (int z = 0; z < ZC; z++) red[w][z][lane] = acc[z];
        __syncthreads();
        for (int z = w; z < ZC; z += 4) {
            T v = 0;
            if (cok && z < zc)
                v = X[t.k1 + col + (long long)(z0 + z) * n] -
                    ((red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]));
            vs[z][lane] = v;
        }
        __syncthreads();
        // x_t = inv(L_tt)' v: column = lane, wave w sums the rows 16w .. 16w+15
#pragma unroll
        for (int z = 0; z < ZC; z++) {
            T a = 0;
#pragma unroll
            for (int j = 0; j < 16; j++) a += iv[j] * vs[z][16 * w + j];
            red[w][z][lane] = a;
        }
        __syncthreads();
        for (int z = w; z < zc; z += 4)
            if (cok)
                st_dev(X + t.k1 + col + (long long)(z0 + z) * n,
                       (red[0][z][lane] + red[1][z][lane]) + (red[2][z][lane] + red[3][z][lane]));
        __syncthreads();
    }
    if (below) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) __hip_atomic_fetch_add(cnt + t.p0, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        chain_publish(cnt + t.p0, nblk - tb);
    }
}

// ---------------------------------------------------------------------------
// FP64 residual R = B - A X with A held as full CSR (both triangles), values
// gathered from tril(A) by src.
// ---------------------------------------------------------------------------
__global__ void k_gather_vals(const double *__restrict__ Ax, const long long *__restrict__ src,
                              long long nnz, double *__restrict__ vals) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; p < nnz; p += stride) vals[p] = Ax[src[p]];
}

template <int LPR>
__global__ __launch_bounds__(256) void k_residual(const long long *__restrict__ rowptr,
                                                  const int *__restrict__ col,
                                                  const double *__restrict__ vals,
                                                  const double *__restrict__ B,
                                                  const double *__restrict__ X,
                                                  double *__restrict__ R, long long n, int nrhs) {
    // LPR lanes per row (4, 8 or 16, from the average row length): the row's
    // values and columns are read coalesced, four right-hand sides per pass
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / LPR;
    const int sub = threadIdx.x & (LPR - 1);
    const bool ok = i < n;
    long long p0 = 0, p1 = 0;
    if (ok) {
        p0 = rowptr[i];
        p1 = rowptr[i + 1];
    }
    for (int z0 = 0; z0 < nrhs; z0 += 4) {
        const int zc = min(4, nrhs - z0);
        double s[4] = {0, 0, 0, 0};
        for (long long p = p0 + sub; p < p1; p += LPR) {
            const double v = vals[p];
            const double *x = X + col[p] + (long long)z0 * n;
#pragma unroll
            for (int z = 0; z < 4; z++)
                if (z < zc) s[z] += v * x[(long long)z * n];
        }
#pragma unroll
        for (int z = 0; z < 4; z++) {
            double u = s[z];
#pragma unroll
            for (int m = LPR / 2; m >= 1; m >>= 1) u += __shfl_xor(u, m);
            if (ok && sub == 0 && z < zc) {
                const long long q = i + (long long)(z0 + z) * n;
                R[q] = B[q] - u;
            }
        }
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
