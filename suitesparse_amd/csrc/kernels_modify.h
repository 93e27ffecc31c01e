// kernels_modify.h -- rank-k update/downdate of the resident supernodal LL'
// factor (CHOLMOD's cholmod_updown, CHOLMOD/Modify/cholmod_updown.c:74):
//   Lnew Lnew' = L L' + C C' (update) or L L' - C C' (downdate),
// C's rows in the factor's order, at most UD_RANK columns per pass.
//
// Each column j on the paths of C's columns is modified by the recurrence of
// Modify/t_cholmod_updown_numkr.c on its LDL' form:
//   pivot (:294-324): d = l_jj^2; per rank r whose path holds j, in order:
//     a = alpha_r +- w_r(j)^2 / d;  d *= a;  alpha_r = a;
//     gamma_r = -+ w_r(j) / d;      d /= alpha_r (old)
//   rows i below j (:363-372): l = l_ij / l_jj (old);
//     per rank: w_r(i) -= w_r(j) l;  l -= gamma_r w_r(i)
//   then l_jj = sqrt(d), l_ij = l * l_jj.
// A 64-column block of a supernode is one workgroup (k_updown_block): lane c
// forms column c's pivot data while the rows of the diagonal block follow one
// column behind.  The rows below the block take the block's 64 columns of pivot
// data in a second launch (k_updown_rows), a lane per row.  The modified
// blocks' inverses (the solve sweeps use them) are formed once at the end, all
// blocks in parallel (k_updown_inv).  Every operation is
// written out and contraction is off, so the result does not depend on the
// block a column falls in or on the launch that takes its rows; a column that
// no path reaches is not touched (its LDL' round trip would not be exact).
#pragma once
#include <hip/hip_runtime.h>
#include "mxc_kernels.h"

namespace mxc {

constexpr int UD_RANK = 8;                 // columns of C per pass (Common->maxrank)
constexpr int UD_PAR = 2 + 2 * UD_RANK;    // per column: old pivot, new pivot, then w_r(j) and gamma_r

// One 64-column block [c0, c0 + nb) of a supernode s on the paths.
struct UdBlock {
    long long lx;            // W offset of supernode s (Lpx[s])
    long long winv;          // W offset of this block's inverse (ld nb)
    int nsrow, nscol;
    int k1, psi;             // first column, first row index position in Ls
    int c0, nb;
    int start[UD_RANK];      // per rank: first column of s on its path, or INT_MAX
    int slot;                // pivot-data slot of this block in its launch
    int pad;
};

// Rows [r0, min(r0 + 64, nsrow)) of the supernode of block blk, below the block.
struct UdRows {
    int blk;
    int r0;
};

// C's entries: W[row * UD_RANK + rank] = value (the host keeps the last of
// repeated entries of a column).
template <typename T>
__global__ __launch_bounds__(256) void k_updown_scatter(const long long *__restrict__ idx,
                                                        const double *__restrict__ val, long long nz,
                                                        T *__restrict__ Wc) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nz) Wc[idx[k]] = (T)val[k];
}

template <typename T>
__global__ __launch_bounds__(64) void k_updown_block(const UdBlock *__restrict__ blocks, T *__restrict__ W,
                                                     T *__restrict__ Wc, T *__restrict__ alpha,
                                                     T *__restrict__ par, int update, double dbound,
                                                     int *__restrict__ flag) {
#pragma clang fp contract(off)
    __shared__ T S[NB][DS];                // the diagonal block, S[i][c]
    __shared__ T P[NB][UD_PAR];
    __shared__ T al[UD_RANK];
    const UdBlock b = blocks[blockIdx.x];
    const int lane = threadIdx.x, nb = b.nb;
    const long long base = b.lx + (long long)b.c0 * b.nsrow + b.c0;   // element (c0, c0)
    {
        T v[NB];                           // every load of the row in flight, then the LDS stores
#pragma unroll
        for (int c = 0; c < NB; c++)
            v[c] = (lane < nb && c <= lane) ? W[base + (long long)c * b.nsrow + lane] : (T)0;
#pragma unroll
        for (int c = 0; c < NB; c++) S[lane][c] = v[c];
    }
    T w[UD_RANK];
    const long long g = (long long)(b.k1 + b.c0 + lane) * UD_RANK;
#pragma unroll
    for (int r = 0; r < UD_RANK; r++) w[r] = lane < nb ? Wc[g + r] : (T)0;
    const int last = b.c0 + nb - 1;
    const bool own = lane < UD_RANK && blocks[blockIdx.x].start[lane] <= last;   // rank `lane` passes here
    if (own) al[lane] = alpha[lane];
    int first = INT_MAX;                   // columns before every path are left alone
#pragma unroll
    for (int r = 0; r < UD_RANK; r++) first = min(first, b.start[r]);
    __syncthreads();

    for (int c = 0; c < nb; c++) {
        const int jl = b.c0 + c;
        if (jl < first) continue;
        if (lane == c) {
            const T lold = S[c][c];
            T dj = lold * lold;
#pragma unroll
            for (int r = 0; r < UD_RANK; r++) {
                if (b.start[r] > jl) continue;
                const T wr = w[r], a0 = al[r];
                T t = wr * wr;
                t = t / dj;
                const T a = update ? a0 + t : a0 - t;
                dj = dj * a;
                al[r] = a;
                P[c][2 + UD_RANK + r] = update ? (-wr) / dj : wr / dj;
                dj = dj / a0;
                P[c][2 + r] = wr;
            }
            if (dbound > 0 && !(dj != dj)) {
                const T bnd = (T)dbound;
                if (dj >= (T)0 ? dj < bnd : dj > -bnd) {
                    dj = dj >= (T)0 ? bnd : -bnd;
                    atomicAdd(flag + 1, 1);
                }
            }
            if (!(dj > (T)0)) atomicMin(flag, b.k1 + jl);
            const T lnew = sqrt(dj);
            S[c][c] = lnew;
            P[c][0] = lold;
            P[c][1] = lnew;
        }
        __syncthreads();
        if (lane > c && lane < nb) {
            T l = S[lane][c] / P[c][0];
#pragma unroll
            for (int r = 0; r < UD_RANK; r++) {
                if (b.start[r] > jl) continue;
                w[r] = w[r] - P[c][2 + r] * l;
                l = l - P[c][2 + UD_RANK + r] * w[r];
            }
            S[lane][c] = l * P[c][1];
        }
    }
    __syncthreads();
    for (int c = 0; c < nb; c++)
        if (lane < nb && lane >= c) W[base + (long long)c * b.nsrow + lane] = S[lane][c];
    T *ps = par + (long long)b.slot * NB * UD_PAR;
    for (int e = lane; e < nb * UD_PAR; e += 64) ps[e] = P[e / UD_PAR][e % UD_PAR];
    if (own) alpha[lane] = al[lane];
}

// The new inverses of the modified diagonal blocks (the solve sweeps use
// them), all at once after the last pass: a workgroup per block; lane c forms
// column c of inv(L) by forward substitution, row after row,
//   x_i = (delta_ic - sum_{c <= k < i} L_ik x_k) / L_ii,
// with four partial sums.
template <typename T>
__global__ __launch_bounds__(64) void k_updown_inv(const DiagTask *__restrict__ tasks, T *__restrict__ W) {
    __shared__ T S[NB][DS];
    __shared__ T X[NB][DS];
    __shared__ T rd[NB];
    const DiagTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x, nb = t.nb;
    {
        T v[NB];
#pragma unroll
        for (int c = 0; c < NB; c++) v[c] = (lane < nb && c <= lane) ? W[t.off + (long long)c * t.ld + lane] : (T)0;
#pragma unroll
        for (int c = 0; c < NB; c++) S[lane][c] = v[c];
    }
    if (lane < nb) rd[lane] = (T)1 / W[t.off + (long long)lane * t.ld + lane];
    __syncthreads();
    for (int i = 0; i < nb; i++) {
        T s0 = (T)0, s1 = (T)0, s2 = (T)0, s3 = (T)0;
        int k = lane;
        for (; k + 3 < i; k += 4) {
            s0 += S[i][k] * X[k][lane];
            s1 += S[i][k + 1] * X[k + 1][lane];
            s2 += S[i][k + 2] * X[k + 2][lane];
            s3 += S[i][k + 3] * X[k + 3][lane];
        }
        for (; k < i; k++) s0 += S[i][k] * X[k][lane];
        const T v = ((i == lane ? (T)1 : (T)0) - ((s0 + s1) + (s2 + s3))) * rd[i];
        X[i][lane] = lane <= i ? v : (T)0;
    }
    __syncthreads();
    for (int c = 0; c < nb; c++)
        if (lane < nb) W[t.winv + lane + (long long)c * nb] = X[lane][c];
}

template <typename T>
__global__ __launch_bounds__(64) void k_updown_rows(const UdBlock *__restrict__ blocks,
                                                    const UdRows *__restrict__ rows, T *__restrict__ W,
                                                    const int *__restrict__ Ls, T *__restrict__ Wc,
                                                    const T *__restrict__ par) {
#pragma clang fp contract(off)
    __shared__ T P[NB * UD_PAR];
    const UdRows rr = rows[blockIdx.x];
    const UdBlock b = blocks[rr.blk];
    const int lane = threadIdx.x, nb = b.nb;
    const T *ps = par + (long long)b.slot * NB * UD_PAR;
    {
        T pv[UD_PAR];
#pragma unroll
        for (int q = 0; q < UD_PAR; q++) pv[q] = lane + 64 * q < nb * UD_PAR ? ps[lane + 64 * q] : (T)0;
#pragma unroll
        for (int q = 0; q < UD_PAR; q++) P[lane + 64 * q] = pv[q];
    }
    int first = INT_MAX;                   // columns before every path are left alone
#pragma unroll
    for (int r = 0; r < UD_RANK; r++) first = min(first, b.start[r]);
    const int ii = rr.r0 + lane;
    const bool on = ii < b.nsrow;
    const long long g = on ? (long long)(ii < b.nscol ? b.k1 + ii : Ls[b.psi + ii]) * UD_RANK : 0;
    T w[UD_RANK];
#pragma unroll
    for (int r = 0; r < UD_RANK; r++) w[r] = on ? Wc[g + r] : (T)0;
    __syncthreads();
    if (!on) return;
    // the row's nb entries and their divisions by the old pivots first (they do
    // not depend on w), then the chain through w, column after column
    T *col = W + b.lx + (long long)b.c0 * b.nsrow + ii;
    T v[NB];
#pragma unroll
    for (int c = 0; c < NB; c++) v[c] = c < nb ? col[(long long)c * b.nsrow] : (T)0;
#pragma unroll
    for (int c = 0; c < NB; c++)
        if (c < nb && b.c0 + c >= first) v[c] = v[c] / P[c * UD_PAR];
#pragma unroll
    for (int c = 0; c < NB; c++) {
        const int jl = b.c0 + c;
        if (c >= nb || jl < first) continue;
        const T *pc = P + c * UD_PAR;
        T l = v[c];
#pragma unroll
        for (int r = 0; r < UD_RANK; r++) {
            if (b.start[r] > jl) continue;
            w[r] = w[r] - pc[2 + r] * l;
            l = l - pc[2 + UD_RANK + r] * w[r];
        }
        v[c] = l * pc[1];
    }
#pragma unroll
    for (int c = 0; c < NB; c++)
        if (c < nb && b.c0 + c >= first) col[(long long)c * b.nsrow] = v[c];
    const int last = b.c0 + nb - 1;
#pragma unroll
    for (int r = 0; r < UD_RANK; r++)
        if (b.start[r] <= last) Wc[g + r] = w[r];
}

}  // namespace mxc
