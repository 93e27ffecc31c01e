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
                           const double *__restrict__ beta, T *__restrict__ W,
                           double *__restrict__ adiag) {
    const double b = *beta;
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    T v = W[dpos[k]];
    if (b != 0.0) {
        v += (T)b;
        W[dpos[k]] = v;
    }
    adiag[k] = (double)v;           // assembled diagonal, for the pivot-ratio test
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// Min / max of the diagonal of L (rcond diagnostic): each block writes
// (min, max, has_nan) of its share to out[3*blockIdx.x ...].
template <typename T>
__global__ void k_diag_minmax(const T *__restrict__ W, const long long *__restrict__ dpos, long long n,
                              double *__restrict__ out) {
    __shared__ double smin[256], smax[256];
    __shared__ int snan;
    if (threadIdx.x == 0) snan = 0;
    __syncthreads();
    double lo = INFINITY, hi = -INFINITY;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const double v = (double)W[dpos[k]];
        if (isnan(v)) snan = 1;
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x] = smin[0];
        out[3 * blockIdx.x + 1] = smax[0];
        out[3 * blockIdx.x + 2] = snan;
    }
}

// ---------------------------------------------------------------------------
// Dense TS x TS update tile on MFMA (see UpdDesc).  TS = 32: one wave owns the
// tile; TS = 64 / 128: four waves, each owning a (TS/2) x (TS/2) quarter, as
// (TS/32)^2 ... i.e. 2x2 or 4x4 blocks of 16x16 MFMA.  The larger tiles do more
// MFMA work per staged byte, which hides the global-load latency of the next
// K chunk (prefetched into registers while the current chunk computes from
// the other LDS buffer).  The MFMA is fed X (C-column side) as its A operand
// and Y (C-row side) as its B operand, so 16 consecutive lanes hold 16
// consecutive C rows: stores into the column-major target are contiguous runs.
// ---------------------------------------------------------------------------
// Workgroup b of a launch is dispatched to XCD b % 8.  Tiles are dealt to
// the XCDs in chunks of 8 consecutive list entries instead (round robin over
// chunks), so the tiles of a chunk, which share an operand strip, run on one
// XCD and share its L2, while the chunks still spread evenly over the XCDs.
// The tail that does not fill a round of 8 chunks keeps its own order.
__device__ __forceinline__ int xcd_tile(int b, int g) {
    constexpr int C = 8;
    const int full = g / (8 * C) * (8 * C);
    if (b >= full) return b;
    const int x = b & 7, k = b >> 3;
    return (x + 8 * (k / C)) * C + (k % C);
}

// Split-FP16 form of the FP32 factor (see k_tile_h): the scale exponent of
// global row g, s_g = 13 - floor(log2 sqrt(a_gg)), and the packed pair
// (hi | lo << 16) of a value v of that row, hi = fp16(v 2^s), lo = fp16(v 2^s - hi).
__device__ __forceinline__ int row_scale_exp(const double *__restrict__ adiag, int g) {
    const double a = adiag[g];
    if (!(a > 0.0) || !(a < 1e300)) return 0;
    return min(max(13 - ilogb(sqrt(a)), -100), 100);
}

__device__ __forceinline__ unsigned split_pack(float v, int e) {
    const float sv = v * ldexpf(1.0f, e);
    const _Float16 hi = (_Float16)sv;
    const _Float16 lo = (_Float16)(sv - (float)hi);
    return (unsigned)__builtin_bit_cast(unsigned short, hi) |
           ((unsigned)__builtin_bit_cast(unsigned short, lo) << 16);
}

// FP64 redo of the diagonal blocks of a low-precision factor whose pivot ratio
// signals breakdown (a pivot at or below rho * a_jj).  cnt: [0] blocks queued
// since the last k_potrf_redo64 launch, [1] blocks refactored in FP64 in this
// factorization, [2] redo workgroups finished (queue clearing).  The rest
// finds A's own entries of a block: the full CSR of A in the original order,
// its values (refreshed every factorization), Perm, the inverse permutation
// and beta.
struct Redo {
    int *cnt;
    DiagTask *list;
    const long long *rowptr;
    const int *col;
    const double *vals;
    const long long *perm;
    const int *pinv;
    const double *beta;
};

__device__ __forceinline__ void redo_push(const Redo &rd, const DiagTask &t) {
    const int k = atomicAdd(rd.cnt, 1);
    rd.list[k] = t;
}

// Refactor diagonal block t in FP64 with the nt threads tid < nt of the
// caller (sync() is their barrier), in place of its low-precision result:
//   1. A64 = the block as assembled in W (A's values in T minus the descendant
//      and earlier-step updates accumulated in T), in FP64, with A's own
//      entries swapped for their FP64 values (+ beta on the diagonal):
//      A64 = W + (A - fl_T(A)).  Column Perm[col0 + c] of A's CSR holds, among
//      its neighbours, the rows whose permuted index falls in the block;
//   2. right-looking Cholesky of A64 in FP64 (a pivot <= 0 is reported as the
//      failing column, as CHOLMOD reports it), L written to W in T;
//   3. inv(L) in place, last column first (column j from the columns right
//      of it: X_ij = -X_jj sum_{k=j+1..i} X_ik L_kj), written to W in T.
// A: NB x (NB + 1) doubles, tmp: NB doubles, sbad: one int (all in LDS).
template <typename T, typename Sync>
__device__ __forceinline__ void redo_block64(const DiagTask &t, T *__restrict__ W, double *A, double *tmp,
                                             int *sbad, int tid, int nt, const Redo &rd,
                                             int *__restrict__ fail, Sync sync) {
    constexpr int LD = NB + 1;
    const int nb = t.nb;
    for (int q = tid; q < nb * nb; q += nt) {
        const int i = q % nb, c = q / nb;
        A[i * LD + c] = (c <= i) ? (double)W[t.off + i + (long long)c * t.ld] : 0.0;
    }
    if (tid == 0) *sbad = -1;
    sync();
    const double b = *rd.beta;
    const int per = nt >= 4 * nb ? 4 : 1;           // threads per column (distinct rows each)
    for (int c = tid / per; c < nb; c += nt / per) {
        const long long o = rd.perm[t.col0 + c];
        for (long long q = rd.rowptr[o] + (tid % per); q < rd.rowptr[o + 1]; q += per) {
            const int i = rd.pinv[rd.col[q]] - t.col0;
            if (i < c || i >= nb) continue;
            const double a = rd.vals[q];
            T a_t = (T)a;                           // what the assembly stored (k_scatter_a, k_add_diag)
            double a64 = a;
            if (i == c && b != 0.0) {
                a_t += (T)b;
                a64 += b;
            }
            A[i * LD + c] += a64 - (double)a_t;
        }
    }
    sync();
    for (int j = 0; j < nb; j++) {
        double d = A[j * LD + j];
        if (!(d > 0.0)) {
            if (tid == 0 && *sbad < 0) *sbad = j;
            d = 1.0;
        }
        const double r = 1.0 / sqrt(d);
        sync();                                     // every thread has read the pivot
        if (tid == 0) A[j * LD + j] = d * r;
        for (int i = j + 1 + tid; i < nb; i += nt) A[i * LD + j] *= r;
        sync();
        const int m = nb - j - 1;
        for (int q = tid; q < m * m; q += nt) {
            const int i = j + 1 + q % m, c = j + 1 + q / m;
            if (c <= i) A[i * LD + c] -= A[i * LD + j] * A[c * LD + j];
        }
        sync();
    }
    for (int q = tid; q < nb * nb; q += nt) {
        const int i = q % nb, c = q / nb;
        if (c <= i) W[t.off + i + (long long)c * t.ld] = (T)A[i * LD + c];
    }
    sync();
    for (int j = nb - 1; j >= 0; j--) {
        if (tid == 0) A[j * LD + j] = 1.0 / A[j * LD + j];
        for (int i = j + 1 + tid; i < nb; i += nt) tmp[i] = A[i * LD + j];
        sync();
        const double xjj = -A[j * LD + j];
        for (int i = j + 1 + tid; i < nb; i += nt) {
            double acc = 0.0;
            for (int k = j + 1; k <= i; k++) acc += A[i * LD + k] * tmp[k];
            A[i * LD + j] = xjj * acc;
        }
        sync();
    }
    for (int q = tid; q < nb * nb; q += nt) {
        const int i = q % nb, c = q / nb;
        W[t.winv + i + (long long)c * nb] = (c <= i) ? (T)A[i * LD + c] : (T)0;
    }
    if (tid == 0) {
        if (*sbad >= 0) atomicMin(fail, t.col0 + *sbad);
        atomicAdd(rd.cnt + 1, 1);
    }
}

template <typename T, bool SAFE>
__device__ __forceinline__ void potrf_block(const DiagTask &t, T *__restrict__ W, T *S, T *X, int *sfail,
                                            int *__restrict__ fail, const double *__restrict__ adiag,
                                            double rho, const Redo &rd, double *tmp);

// Arguments of the diagonal-block tail of a trailing launch (PD): after its
// epilogue, the workgroup of tile (0, 0) of a block with pd >= 0 has the final
// values of the next step's diagonal block and factors it (potrf_block).
//
// With a folded TRSM (UpdDesc::fold) the tail publishes the factored block:
// ready[pd] is set once its L and inverse are in W (a device-scope release
// after the workgroup's stores), and the tiles of column 0 below it wait for
// that flag (device-scope acquire) before applying the inverse.  Producer
// tiles come first in the launch (the first nfirst workgroups map to tiles
// 0 .. nfirst-1, which hold them) and never wait, and workgroups start in
// index order, so every wait ends; a wait is still bounded (FOLD_SPIN polls),
// past which it raises *abort and the host reports a device error.
struct PdArgs {
    const DiagTask *diags;
    int *fail;
    double rho;
    Redo rd;
    int *ready;        // per diagonal task: 1 once a folding tail has published it
    int *abort;        // set by a wait that gave up
    int nfirst;        // producer tiles at the head of the launch (no XCD remap)
};

constexpr int FOLD_SPIN = 1 << 22;

// The tail itself: S and X of potrf_block carved from the tile's LDS buffer.
template <typename T>
__device__ __forceinline__ void pd_tail(const PdArgs &pda, int pd, T *__restrict__ W, unsigned char *smem,
                                        const double *__restrict__ adiag, bool publish) {
    __shared__ int sfail[2];
    __syncthreads();                        // the tile's stores and LDS reads are done
    T *S = reinterpret_cast<T *>(smem);
    if (pda.rho > 0) {
        __shared__ double tmp[NB];
        potrf_block<T, true>(pda.diags[pd], W, S, S + NB * DS, sfail, pda.fail, adiag, pda.rho, pda.rd, tmp);
    } else {
        potrf_block<T, false>(pda.diags[pd], W, S, S + NB * DS, sfail, pda.fail, adiag, pda.rho, pda.rd, nullptr);
    }
    if (publish) {
        __syncthreads();                    // every wave's stores of L and the inverse are issued
        if (threadIdx.x == 0) __hip_atomic_store(pda.ready + pd, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Wait (whole workgroup) until *flag != 0 or the wait gives up.
__device__ __forceinline__ void fold_wait(const int *flag, int *abort) {
    if ((threadIdx.x & 63) == 0) {
        for (int it = 1; __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0; it++) {
            __builtin_amdgcn_s_sleep(1);
            if ((it & 63) == 0 &&
                (it >= FOLD_SPIN || __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the inverse's loads come after the flag
    __syncthreads();
}

// One tile of a launch: virtual workgroup bx of g (k_tile: the workgroup
// itself; k_tile_grid: each of its workgroups takes bx, bx + grid, ...).
template <typename T, int MODE, int TS, bool PD>
__device__ __forceinline__ void tile_work(const UpdDesc *__restrict__ descs, const Tile *__restrict__ tiles,
                                          T *__restrict__ W, const int *__restrict__ Ls, unsigned *__restrict__ Lh,
                                          const double *__restrict__ adiag, const PdArgs &pda, const int bx,
                                          const int g) {
    typedef typename MfmaOp<T>::acc_t acc_t;
    constexpr int NT = TS <= 32 ? 64 : 256;     // threads (tiles of 16 and 32: one wave)
    constexpr int WB = TS <= 32 ? TS : TS / 2;  // edge of a wave's block of C
    constexpr int MB = WB / 16;                 // 16x16 MFMA blocks per wave edge
    constexpr int KG = NT / TS;                 // k rows staged per pass
    constexpr int PER = (BK * TS) / NT;         // elements of X (and of Y) each thread stages
    static_assert(!PD || (MODE == MODE_TRAIL && TS >= 64), "diagonal tails: trailing tiles of 64 or 128");
    // one LDS buffer: the staged X / Y chunks, reused by the diagonal-block tail.
    // Tiles of 32 and 64 stage into a single buffer: the extra barrier per
    // chunk costs less than the resident tiles that half the LDS adds (64³ FP64
    // update kernel 28.1 -> 30.2 TFLOP/s); a 128 tile is register-bound at two
    // waves per SIMD either way and keeps two buffers.
    constexpr int NBUF = TS == 128 ? 2 : 1;
    constexpr size_t XY_BYTES = sizeof(T) * NBUF * BK * (TS + LDS_PAD);
    constexpr size_t PD_BYTES = PD ? sizeof(T) * 2 * NB * DS : 0;
    // folded TRSM (FP64, 64-wide trailing tiles): the updated block and the
    // next diagonal block's inverse, both k-major like the staged chunks
    constexpr bool CAN_FOLD = PD && MODE == MODE_TRAIL && TS == NB && std::is_same<T, double>::value;
    constexpr size_t FOLD_HALF = sizeof(T) * NB * (TS + LDS_PAD);
    constexpr size_t FOLD_BYTES = CAN_FOLD ? 2 * FOLD_HALF : 0;
    constexpr size_t SMEM0 = 2 * XY_BYTES > PD_BYTES ? 2 * XY_BYTES : PD_BYTES;
    constexpr size_t SMEM = SMEM0 > FOLD_BYTES ? SMEM0 : FOLD_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];
    T (*Xs)[BK][TS + LDS_PAD] = reinterpret_cast<T (*)[BK][TS + LDS_PAD]>(smem);
    T (*Ys)[BK][TS + LDS_PAD] = reinterpret_cast<T (*)[BK][TS + LDS_PAD]>(smem + XY_BYTES);
    __shared__ int relr[TS];
    __shared__ int relc[TS];

    const int nf = PD ? pda.nfirst : 0;
    const Tile tl = tiles[bx < nf ? bx : nf + xcd_tile(bx - nf, g - nf)];
    const UpdDesc d = descs[tl.desc];
    const int i0 = tl.ti * TS, j0 = tl.tj * TS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = TS <= 32 ? 0 : (wave & 1), wc = TS <= 32 ? 0 : (wave >> 1);

    if (MODE == MODE_SCATTER) {
        for (int e = tid; e < 2 * TS; e += NT) {
            if (e < TS) {
                int i = i0 + e;
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
                relr[e] = r;
            } else {
                int j = j0 + e - TS;
                relc[e - TS] = (j < d.N) ? (Ls[d.ls_off + j] - d.base) : -1;
            }
        }
    }

    acc_t acc[MB][MB];
#pragma unroll
    for (int a = 0; a < MB; a++)
#pragma unroll
        for (int b = 0; b < MB; b++) acc[a][b] = acc_t{0, 0, 0, 0};

    const T *__restrict__ Xp = W + d.x_off;
    const T *__restrict__ Yp = W + d.y_off;
    // thread t stages row r = t % TS of chunk rows k = t/TS + KG*q, q < PER
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
        if (NBUF == 1 && k0 > 0) __syncthreads();   // the previous chunk's reads are done
#pragma unroll
        for (int q = 0; q < PER; q++) {
            Ys[buf][kq + KG * q][r] = ry[q];
            Xs[buf][kq + KG * q][r] = rx[q];
        }
        __syncthreads();
        if (k0 + BK < d.K) fetch(k0 + BK);      // next chunk in flight during the MFMAs
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T xa[MB], yb[MB];
#pragma unroll
            for (int m = 0; m < MB; m++) xa[m] = Xs[buf][kk + (lane >> 4)][wc * WB + m * 16 + (lane & 15)];
#pragma unroll
            for (int n = 0; n < MB; n++) yb[n] = Ys[buf][kk + (lane >> 4)][wr * WB + n * 16 + (lane & 15)];
#pragma unroll
            for (int m = 0; m < MB; m++)
#pragma unroll
                for (int n = 0; n < MB; n++) acc[m][n] = MfmaOp<T>::mma(xa[m], yb[n], acc[m][n]);
        }
        if constexpr (NBUF == 2) buf ^= 1;
    }
    if (MODE == MODE_STORE) __syncthreads();   // in-place TRSM: every read of the panel is done
    const bool fold = CAN_FOLD && d.fold && d.pd >= 0 && tl.tj == 0 && tl.ti >= 1;
    T (*Af)[TS + LDS_PAD] = reinterpret_cast<T (*)[TS + LDS_PAD]>(smem);               // Af[k][i]
    T (*If)[TS + LDS_PAD] = reinterpret_cast<T (*)[TS + LDS_PAD]>(smem + FOLD_HALF);   // If[k][j]
    if (CAN_FOLD && fold) __syncthreads();     // every wave is done with the staged chunks

    // FP32 TRSM with Lh: the panel is also stored in split-FP16 form, each row
    // scaled by its own exponent, for the update and trailing tiles that read it
    int exs[MB];
#pragma unroll
    for (int n = 0; n < MB; n++) exs[n] = 0;
    if (MODE == MODE_STORE && std::is_same<T, float>::value && Lh) {
#pragma unroll
        for (int n = 0; n < MB; n++) {
            const int i = i0 + wr * WB + n * 16 + (lane & 15);
            if (i < d.M) exs[n] = row_scale_exp(adiag, Ls[d.ls_off + i]);
        }
    }

    // epilogue: acc[m][n][v] = C(i, j), i = i0 + WB*wr + 16n + lane%16,
    //                                   j = j0 + WB*wc + 16m + row(lane, v)
#pragma unroll
    for (int m = 0; m < MB; m++) {
#pragma unroll
        for (int n = 0; n < MB; n++) {
            const int il = wr * WB + n * 16 + (lane & 15);
            const int i = i0 + il;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int jl = wc * WB + m * 16 + MfmaOp<T>::row(lane, v);
                const int j = j0 + jl;
                if (i >= d.M || j >= d.N) continue;
                const T c = acc[m][n][v];
                if (MODE == MODE_SCATTER) {
                    if (i < j) continue;
                    atomic_sub(W + d.c_off + (long long)relc[jl] * d.ldc + relr[il], c);
                } else if (MODE == MODE_TRAIL) {
                    if (i < j) continue;
                    T *w = W + d.c_off + (long long)j * d.ldc + i;
                    if (CAN_FOLD && fold) Af[jl][il] = *w - c;   // the next TRSM's input, kept here
                    else *w -= c;
                } else {
                    W[d.c_off + (long long)j * d.ldc + i] = c;
                    if (std::is_same<T, float>::value && Lh)
                        Lh[d.c_off + (long long)j * d.ldc + i] = split_pack((float)c, exs[n]);
                }
            }
        }
    }
    if constexpr (CAN_FOLD) {
        if (fold) {
            // the next step's TRSM of these rows: L(i, j) = sum_k A(i, k) inv(j, k),
            // with the MFMA sequence of the TRSM tile (k ascending)
            const DiagTask dt = pda.diags[d.pd];
            fold_wait(pda.ready + d.pd, pda.abort);
            for (int e = tid; e < NB * NB; e += NT) If[e / NB][e % NB] = W[dt.winv + e];
            __syncthreads();
            acc_t a2[MB][MB];
#pragma unroll
            for (int a = 0; a < MB; a++)
#pragma unroll
                for (int b = 0; b < MB; b++) a2[a][b] = acc_t{0, 0, 0, 0};
#pragma unroll 4
            for (int kk = 0; kk < NB; kk += 4) {
                T xa[MB], yb[MB];
#pragma unroll
                for (int m = 0; m < MB; m++) xa[m] = If[kk + (lane >> 4)][wc * WB + m * 16 + (lane & 15)];
#pragma unroll
                for (int n = 0; n < MB; n++) yb[n] = Af[kk + (lane >> 4)][wr * WB + n * 16 + (lane & 15)];
#pragma unroll
                for (int m = 0; m < MB; m++)
#pragma unroll
                    for (int n = 0; n < MB; n++) a2[m][n] = MfmaOp<T>::mma(xa[m], yb[n], a2[m][n]);
            }
#pragma unroll
            for (int m = 0; m < MB; m++) {
#pragma unroll
                for (int n = 0; n < MB; n++) {
                    const int i = i0 + wr * WB + n * 16 + (lane & 15);
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int j = wc * WB + m * 16 + MfmaOp<T>::row(lane, v);
                        if (i < d.M) W[d.c_off + (long long)j * d.ldc + i] = a2[m][n][v];
                    }
                }
            }
        }
    }
    if constexpr (PD) {
        if (d.pd >= 0 && tl.ti == 0 && tl.tj == 0) pd_tail<T>(pda, d.pd, W, smem, adiag, d.fold != 0);
    }
}

template <typename T, int MODE, int TS, bool PD = false>
__global__ __launch_bounds__(TS <= 32 ? 64 : 256, TS == 128 ? 2 : 1) void k_tile(const UpdDesc *__restrict__ descs,
                                                              const Tile *__restrict__ tiles,
                                                              T *__restrict__ W,
                                                              const int *__restrict__ Ls,
                                                              unsigned *__restrict__ Lh,
                                                              const double *__restrict__ adiag,
                                                              PdArgs pda = PdArgs{}) {
    tile_work<T, MODE, TS, PD>(descs, tiles, W, Ls, Lh, adiag, pda, (int)blockIdx.x, (int)gridDim.x);
}

// The tiles of a launch on a grid of at most a few workgroups per CU (a
// multiple of 8, so a workgroup's tiles keep its XCD): the second-stream
// trailing update leaves room on every CU for the diagonal factorizations
// of the main stream, which otherwise wait until the whole launch drains.
template <typename T, int MODE, int TS>
__global__ __launch_bounds__(256, 1) void k_tile_grid(const UpdDesc *__restrict__ descs,
                                                      const Tile *__restrict__ tiles, T *__restrict__ W,
                                                      const int *__restrict__ Ls, const double *__restrict__ adiag,
                                                      int ntiles) {
    static_assert(TS >= 64, "four-wave tiles");
    for (int b = (int)blockIdx.x; b < ntiles; b += (int)gridDim.x) {
        tile_work<T, MODE, TS, false>(descs, tiles, W, Ls, nullptr, adiag, PdArgs{}, b, ntiles);
        __syncthreads();                   // the next tile restages the LDS buffers
    }
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// FP32 factor, split-FP16 products (update and trailing blocks).  An FP32
// value v of L in global row g is scaled by 2^s_g, s_g = 13 - floor(log2
// sqrt(a_gg)) with a_gg the assembled diagonal of A (|L_gk| <= sqrt(a_gg), so
// the scaled values stay below 2^14), and split into hi = fp16(v 2^s) and
// lo = fp16(v 2^s - hi).  The TRSM that finalizes a value writes the pair
// once into Lh (same layout as L), so the tiles here stage the halves as they
// are.  C = Y X' accumulates hi*hi + hi*lo + lo*hi in FP32 on the 16x16x32
// FP16 MFMA (lo*lo lies below FP32 rounding relative to the row norms) and is
// unscaled by 2^-(s_i + s_j) in the epilogue, which is the one of k_tile.
// One LDS buffer of hi / lo rows, k contiguous, row stride 40 halves (the
// 16-byte reads of a 16-row group fall in distinct banks); the next chunk is
// prefetched into registers.
// ---------------------------------------------------------------------------
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

template <int MODE, int TS, bool PD = false>
__global__ __launch_bounds__(256, 2) void k_tile_h(const UpdDesc *__restrict__ descs,
                                                                  const Tile *__restrict__ tiles,
                                                                  float *__restrict__ W,
                                                                  const unsigned *__restrict__ Lh,
                                                                  const int *__restrict__ Ls,
                                                                  const double *__restrict__ adiag,
                                                                  PdArgs pda = PdArgs{}) {
    static_assert(TS == 64 || TS == 128, "split-FP16 tiles are 64 or 128 wide");
    constexpr int NT = 256;              // 2 x 2 waves
    constexpr int WB = TS / 2;           // edge of a wave's block of C
    constexpr int MB = WB / 16;
    constexpr int BKH = 32;              // K per chunk: one MFMA depth
    constexpr int RS = BKH + 8;          // LDS row stride (halves)
    constexpr int KG = NT / TS;          // staging threads per row
    constexpr int PER = BKH / KG;        // consecutive k per staging thread
    static_assert(!PD || (MODE == MODE_TRAIL && TS >= 64), "diagonal tails: trailing tiles of 64 or 128");
    // one LDS buffer: the staged halves, reused by the diagonal-block tail
    constexpr size_t H_BYTES = 4 * sizeof(_Float16) * TS * RS;
    constexpr size_t PD_BYTES = PD ? sizeof(float) * 2 * NB * DS : 0;
    __shared__ __attribute__((aligned(16))) unsigned char smem[H_BYTES > PD_BYTES ? H_BYTES : PD_BYTES];
    _Float16 *Xh = reinterpret_cast<_Float16 *>(smem);
    _Float16 *Xl = Xh + TS * RS;
    _Float16 *Yh = Xl + TS * RS;
    _Float16 *Yl = Yh + TS * RS;
    __shared__ int relr[TS], relc[TS];
    __shared__ int exr[TS], exc[TS];     // scale exponents of the Y rows (i) and X rows (j)

    const Tile tl = tiles[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
    const UpdDesc d = descs[tl.desc];
    const int i0 = tl.ti * TS, j0 = tl.tj * TS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;

    for (int e = tid; e < 2 * TS; e += NT) {
        if (e < TS) {
            const int i = i0 + e;
            int r = -1, ex = 0;
            if (i < d.M) {
                const int g = Ls[d.ls_off + i];
                ex = row_scale_exp(adiag, g);
                if (MODE == MODE_SCATTER) {
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
            }
            relr[e] = r;
            exr[e] = ex;
        } else {
            const int j = j0 + e - TS;
            int c = -1, ex = 0;
            if (j < d.N) {
                const int g = Ls[d.ls_off + j];
                ex = row_scale_exp(adiag, g);
                c = g - d.base;
            }
            relc[e - TS] = c;
            exc[e - TS] = ex;
        }
    }
    __syncthreads();

    // staging: thread t holds row r = t % TS, k in [kq * PER, kq * PER + PER) of a chunk
    const int r = tid % TS, kq = tid / TS;
    const bool rowy = (i0 + r) < d.M, rowx = (j0 + r) < d.N;
    const unsigned *yrow = Lh + d.y_off + (i0 + r);
    const unsigned *xrow = Lh + d.x_off + (j0 + r);
    unsigned ry[PER], rx[PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int gk = k0 + kq * PER + q;
            const bool kin = gk < d.K;
            ry[q] = (kin && rowy) ? yrow[(long long)gk * d.ldy] : 0u;
            rx[q] = (kin && rowx) ? xrow[(long long)gk * d.ldx] : 0u;
        }
    };
    auto half_of = [](unsigned u) { return __builtin_bit_cast(_Float16, (unsigned short)(u & 0xffffu)); };
    auto stage = [&]() {
#pragma unroll
        for (int h = 0; h < PER; h += 8) {
            half8_t yh, yl, xh, xl;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                yh[q] = half_of(ry[h + q]);
                yl[q] = half_of(ry[h + q] >> 16);
                xh[q] = half_of(rx[h + q]);
                xl[q] = half_of(rx[h + q] >> 16);
            }
            const int o = r * RS + kq * PER + h;
            *reinterpret_cast<half8_t *>(&Yh[o]) = yh;
            *reinterpret_cast<half8_t *>(&Yl[o]) = yl;
            *reinterpret_cast<half8_t *>(&Xh[o]) = xh;
            *reinterpret_cast<half8_t *>(&Xl[o]) = xl;
        }
    };

    f32x4 acc[MB][MB];
#pragma unroll
    for (int a = 0; a < MB; a++)
#pragma unroll
        for (int b = 0; b < MB; b++) acc[a][b] = f32x4{0, 0, 0, 0};

    fetch(0);
    const int ko = 8 * (lane >> 4);
    for (int k0 = 0; k0 < d.K; k0 += BKH) {
        __syncthreads();                         // the previous chunk's reads are done
        stage();
        __syncthreads();
        if (k0 + BKH < d.K) fetch(k0 + BKH);     // next chunk in flight during the MFMAs
        half8_t xh[MB], xl[MB], yh[MB], yl[MB];
#pragma unroll
        for (int m = 0; m < MB; m++) {
            const int o = (wc * WB + m * 16 + (lane & 15)) * RS + ko;
            xh[m] = *reinterpret_cast<const half8_t *>(&Xh[o]);
            xl[m] = *reinterpret_cast<const half8_t *>(&Xl[o]);
        }
#pragma unroll
        for (int n = 0; n < MB; n++) {
            const int o = (wr * WB + n * 16 + (lane & 15)) * RS + ko;
            yh[n] = *reinterpret_cast<const half8_t *>(&Yh[o]);
            yl[n] = *reinterpret_cast<const half8_t *>(&Yl[o]);
        }
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < MB; n++) {
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh[m], yh[n], acc[m][n], 0, 0, 0);
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh[m], yl[n], acc[m][n], 0, 0, 0);
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl[m], yh[n], acc[m][n], 0, 0, 0);
            }
    }

    // epilogue: acc[m][n][v] = C(i, j) 2^(s_i + s_j), i = i0 + WB*wr + 16n + lane%16,
    //                                               j = j0 + WB*wc + 16m + 4 (lane/16) + v
#pragma unroll
    for (int m = 0; m < MB; m++) {
#pragma unroll
        for (int n = 0; n < MB; n++) {
            const int il = wr * WB + n * 16 + (lane & 15);
            const int i = i0 + il;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int jl = wc * WB + m * 16 + 4 * (lane >> 4) + v;
                const int j = j0 + jl;
                if (i >= d.M || j >= d.N || i < j) continue;
                const float c = ldexpf(acc[m][n][v], -(exr[il] + exc[jl]));
                if (MODE == MODE_SCATTER)
                    atomic_sub(W + d.c_off + (long long)relc[jl] * d.ldc + relr[il], c);
                else
                    W[d.c_off + (long long)j * d.ldc + i] -= c;
            }
        }
    }
    if constexpr (PD) {
        if (d.pd >= 0 && tl.ti == 0 && tl.tj == 0) pd_tail<float>(pda, d.pd, W, smem, adiag, false);
    }
}

// Diagonal block Cholesky + triangular inverse of a 64x64 block (a block
// narrower than 64 is extended by the identity), one 256-thread workgroup,
// block-recursive over 16x16 sub-blocks with the 16x16x4 MFMA:
//   for p = 0..3:  POTRF + inverse of D_pp (one wave, registers)
//                  panel TRSM  L_tp = A_tp inv(D_pp)'      (MFMA, t > p)
//                  trailing    A_ij -= L_ip L_jp'          (MFMA, p < j <= i)
//   inverse:       X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj  (MFMA, block rows i)
// 18 barriers in all.  The lower factor is written back in place (the strict
// upper triangle is left alone, as dpotrf leaves it) and inv(L) (zero above
// the diagonal) goes to the block's inverse slot (ld nb).  A non-positive
// pivot at column j records col0 + j in *fail (atomic min).
// ---------------------------------------------------------------------------

// 1/sqrt(x) for x > 0: the hardware estimate refined by two Newton steps
// y <- y + y (1 - x y^2) / 2 (each roughly doubles the correct bits).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const double e = fma(-x * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}
__device__ __forceinline__ float rsqrt_nr(float x) {
    float y = __builtin_amdgcn_rsqf(x);
    const float e = fmaf(-x * y, y, 1.0f);
    return fmaf(0.5f * y, e, y);
}

// LDS stores of one lane visible to the other lanes of its wave
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of v in lane k (k wave-uniform), through a scalar register
__device__ __forceinline__ double lane_bcast(double v, int k) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_bcast(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// acc += X * Y' for 16x16 blocks X at S[rx..][cx..], Y at S[ry..][cy..] (row-major, stride DS)
template <typename T>
__device__ __forceinline__ void mma_nt16(const T *S, int rx, int cx, const T *S2, int ry, int cy,
                                         typename MfmaOp<T>::acc_t &acc, int lane) {
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
        const T a = S[(rx + (lane & 15)) * DS + cx + kk + (lane >> 4)];
        const T b = S2[(ry + (lane & 15)) * DS + cy + kk + (lane >> 4)];
        acc = MfmaOp<T>::mma(a, b, acc);
    }
}

// acc += X * Y for 16x16 blocks X at S[rx..][cx..], Y at S2[ry..][cy..]
template <typename T>
__device__ __forceinline__ void mma_nn16(const T *S, int rx, int cx, const T *S2, int ry, int cy,
                                         typename MfmaOp<T>::acc_t &acc, int lane) {
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
        const T a = S[(rx + (lane & 15)) * DS + cx + kk + (lane >> 4)];
        const T b = S2[(ry + kk + (lane >> 4)) * DS + cy + (lane & 15)];
        acc = MfmaOp<T>::mma(a, b, acc);
    }
}

// C at S[r0..][c0..] (op) acc, where acc holds D[row(lane,v)][lane%16] = C(m, n) with
// m indexing the MFMA A side (rows) and n the B side (columns)
template <typename T>
__device__ __forceinline__ void store16(T *S, int r0, int c0, const typename MfmaOp<T>::acc_t &acc,
                                        int lane, T alpha, bool accumulate) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
        T *q = S + (r0 + MfmaOp<T>::row(lane, v)) * DS + c0 + (lane & 15);
        *q = accumulate ? *q + alpha * acc[v] : alpha * acc[v];
    }
}

// POTRF + inverse of the 16x16 diagonal block at S[R.., R..] (row stride ld)
// by one wave, lane i < 16 owning row i; the inverse (zero above the diagonal)
// goes to X in the same layout.  Each pivot costs one reciprocal square root
// (no divide on the serial path): L_jj = d_jj r_j and L_ij = a_ij r_j with
// r_j = 1/sqrt(d_jj); the inverse reuses r_j = 1/L_jj.  Returns the first
// non-positive pivot of the block (16 if none), wave-uniform.  SAFE: *brk is
// set when a pivot falls to rho * a_jj or below (a_jj the assembled diagonal
// of A): the block's result is then not used (k_potrf_redo64 refactors it).
//
// The inverse is formed column by column (lane c: x_i = r_i (delta_ic -
// sum_{k<i} L_ik x_k)) with each partial sum updated as soon as x_k is
// final, the same products in the same order as the row-by-row form.
template <typename T, bool SAFE>
__device__ __forceinline__ int potrf16_wave(T *S, T *X, int ld, int R, int nb, int lane,
                                            const double *__restrict__ adiag, int col0, double rho,
                                            bool *brk) {
    T a[16], rinv[16];
#pragma unroll
    for (int c = 0; c < 16; c++) a[c] = (lane < 16) ? S[(R + lane) * ld + R + c] : (T)0;
    // pivot floors of the panel's columns (low-precision safeguard), lane j holds column j's
    T floor_l = 0;
    if (SAFE && lane < 16 && R + lane < nb) floor_l = (T)(rho * adiag[col0 + R + lane]);
    // wave-uniform bookkeeping
    bool low_any = false;   // a pivot at or below its floor
    int badj = 16;          // first non-positive pivot of the panel
#pragma unroll
    for (int j = 0; j < 16; j++) {
        T djj = lane_bcast(a[j], j);
        const bool real = R + j < nb;
        if (SAFE) low_any = low_any || (real && djj <= lane_bcast(floor_l, j));
        const bool bad = !(djj > (T)0);
        badj = (bad && real && badj == 16) ? j : badj;
        djj = bad ? (T)1 : djj;
        const T rj = rsqrt_nr(djj);
        rinv[j] = rj;
        const T l = (lane > j) ? a[j] * rj : ((lane == j) ? djj * rj : (T)0);
        a[j] = l;
#pragma unroll
        for (int c = j + 1; c < 16; c++) a[c] -= l * lane_bcast(l, c);
    }
    if (SAFE && low_any) *brk = true;
    if (lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; c++)
            if (c <= lane) S[(R + lane) * ld + R + c] = a[c];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // inverse: lane c < 16 forms column c
    T x[16], sm[16];
#pragma unroll
    for (int i = 0; i < 16; i++) sm[i] = (i == lane) ? (T)1 : (T)0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        x[k] = sm[k] * rinv[k];
#pragma unroll
        for (int i = k + 1; i < 16; i++) sm[i] -= S[(R + i) * ld + R + k] * x[k];
    }
    if (lane < 16) {
#pragma unroll
        for (int i = 0; i < 16; i++) X[(R + i) * ld + R + lane] = (i >= lane) ? x[i] : (T)0;
    }
    return badj;
}

// SAFE (rho > 0, low-precision factor): a pivot d_jj <= rho * a_jj (a_jj the
// assembled diagonal of A) signals breakdown of the block.  Its low-precision
// result is dropped and the same workgroup refactors the block in FP64 in
// place (redo_block64, in the LDS of S and X, which hold NB x (NB + 1)
// doubles for T = float), before anything reads it.  Without SAFE a pivot
// <= 0 is reported as the failing column.
//
// potrf_block: the whole workgroup (256 threads) factors diagonal block t in
// LDS (S: the block, X: its inverse, both NB x DS row-major and contiguous;
// sfail: two shared ints, the first failing column and the breakdown flag;
// tmp: NB shared doubles for the redo) and writes L_tt and the inverse back
// to W.
template <typename T, bool SAFE>
__device__ __forceinline__ void potrf_block(const DiagTask &t, T *__restrict__ W, T *S, T *X, int *sfail,
                                            int *__restrict__ fail, const double *__restrict__ adiag,
                                            double rho, const Redo &rd, double *tmp) {
    typedef typename MfmaOp<T>::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nb = t.nb;
    const int np = (nb + 15) >> 4;          // 16-column panels that hold real columns
    {
        // all 16 loads of a thread in flight at once, then the LDS stores
        T v[16];
        const int i = tid & (NB - 1), c0 = tid >> 6;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = c0 + 4 * q;
            v[q] = (i < nb && c < nb && c <= i) ? W[t.off + i + (long long)c * t.ld] : (T)0;
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = c0 + 4 * q;
            S[i * DS + c] = (i >= nb && i == c) ? (T)1 : v[q];
            X[i * DS + c] = 0;
        }
    }
    if (tid == 0) {
        sfail[0] = -1;
        sfail[1] = 0;
    }
    __syncthreads();

    // Right-looking over 16-column panels, one panel of look-ahead: after the
    // TRSM of panel p, wave 0 updates only the next diagonal block and factors
    // it at once, while waves 1..3 apply panel p's other updates and form block
    // row p of the inverse, whose inputs are final by then:
    //   L_ii, X_ii = potrf16_wave(S_ii);   L_ip = S_ip X_pp';
    //   S_ij -= L_ip L_jp' (p < j <= i);   X_pj = -X_pp sum_{k=j}^{p-1} L_pk X_kj (j < p).
    if (wave == 0) {
        bool brk = false;
        const int badj = potrf16_wave<T, SAFE>(S, X, DS, 0, nb, lane, adiag, t.col0, rho, &brk);
        if (lane == 0 && badj < 16 && sfail[0] < 0) sfail[0] = badj;
        if (SAFE && lane == 0 && brk) sfail[1] = 1;
    }
    __syncthreads();
    for (int p = 0; p < np; p++) {
        const int R = 16 * p;
        // (1) panel TRSM: rows below, 16-row tiles p+1..np-1, one wave each
        if (wave < np - 1 - p) {
            const int rt = 16 * (p + 1 + wave);
            acc_t acc = acc_t{0, 0, 0, 0};
            // C(m = row of inv, n = panel row) = sum_k X(m, k) A(n, k)  ->  L(n, m)
            mma_nt16<T>(X, R, R, S, rt, R, acc, lane);
#pragma unroll
            for (int v = 0; v < 4; v++)
                S[(rt + (lane & 15)) * DS + R + MfmaOp<T>::row(lane, v)] = acc[v];
        }
        __syncthreads();
        if (wave == 0) {
            // (2a) the next diagonal block: its update by panel p, then its POTRF
            if (p + 1 < np) {
                const int Q = R + 16;
                acc_t acc = acc_t{0, 0, 0, 0};
                mma_nt16<T>(S, Q, R, S, Q, R, acc, lane);
#pragma unroll
                for (int v = 0; v < 4; v++) S[(Q + (lane & 15)) * DS + Q + MfmaOp<T>::row(lane, v)] -= acc[v];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                bool brk = false;
                const int badj = potrf16_wave<T, SAFE>(S, X, DS, Q, nb, lane, adiag, t.col0, rho, &brk);
                if (lane == 0 && badj < 16 && sfail[0] < 0) sfail[0] = Q + badj;
                if (SAFE && lane == 0 && brk) sfail[1] = 1;
            }
        } else {
            // (2b) waves 1..3, round robin: panel p's other updates, then block row p of the inverse
            int q = 0;
            for (int i = p + 1; i < np; i++)
                for (int j = p + 1; j <= i; j++) {
                    if (i == p + 1 && j == p + 1) continue;
                    if ((q++ % 3) != wave - 1) continue;
                    acc_t acc = acc_t{0, 0, 0, 0};
                    // C(m = row in block j, n = row in block i) = sum_k L_jp(m,k) L_ip(n,k)
                    mma_nt16<T>(S, 16 * j, R, S, 16 * i, R, acc, lane);
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        T *dst = S + (16 * i + (lane & 15)) * DS + 16 * j + MfmaOp<T>::row(lane, v);
                        *dst -= acc[v];
                    }
                }
            for (int j = 0; j < p; j++) {
                if ((q++ % 3) != wave - 1) continue;
                acc_t tacc = acc_t{0, 0, 0, 0};
                for (int k = j; k < p; k++) mma_nn16<T>(S, R, 16 * k, X, 16 * k, 16 * j, tacc, lane);
                // stage T = sum_k L_pk X_kj in X's (still zero) block (p, j), then X_pj = -X_pp T
                store16<T>(X, R, 16 * j, tacc, lane, (T)1, false);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                acc_t acc = acc_t{0, 0, 0, 0};
                mma_nn16<T>(X, R, R, X, R, 16 * j, acc, lane);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                store16<T>(X, R, 16 * j, acc, lane, (T)-1, false);
            }
        }
        __syncthreads();
    }
    if (SAFE && sfail[1]) {
        // breakdown: the block again from W, in FP64, in the LDS of S and X
        static_assert(!SAFE || sizeof(T) * 2 * NB * DS >= sizeof(double) * NB * (NB + 1),
                      "the FP64 redo needs the LDS of S and X");
        redo_block64<T>(t, W, reinterpret_cast<double *>(S), tmp, sfail, tid, 256, rd, fail,
                        [] { __syncthreads(); });
        return;
    }
    if (tid == 0 && sfail[0] >= 0) atomicMin(fail, t.col0 + sfail[0]);
    {
        const int i = tid & (NB - 1), c0 = tid >> 6;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = c0 + 4 * q;
            if (i < nb && c < nb) {
                if (c <= i) W[t.off + i + (long long)c * t.ld] = S[i * DS + c];
                W[t.winv + i + (long long)c * nb] = (c <= i) ? X[i * DS + c] : (T)0;
            }
        }
    }
}

template <typename T, bool SAFE>
__global__ __launch_bounds__(256) void k_potrf_diag(const DiagTask *__restrict__ tasks,
                                                    T *__restrict__ W, int *__restrict__ fail,
                                                    const double *__restrict__ adiag, double rho, Redo rd) {
    __shared__ __attribute__((aligned(16))) T SX[2 * NB * DS];   // the block and its inverse, row-major
    __shared__ int sfail[2];
    __shared__ double tmp[SAFE ? NB : 1];
    potrf_block<T, SAFE>(tasks[blockIdx.x], W, SX, SX + NB * DS, sfail, fail, adiag, rho, rd, tmp);
}


// Diagonal blocks of at most 16 columns (the many tiny supernodes of the
// lower etree levels): one wave per block, four blocks per workgroup, the
// same POTRF + inverse as panel 0 of k_potrf_diag.
template <typename T, bool SAFE>
__global__ __launch_bounds__(256) void k_potrf_small(const DiagTask *__restrict__ tasks, int ntask,
                                                     T *__restrict__ W, int *__restrict__ fail,
                                                     const double *__restrict__ adiag, double rho, Redo rd) {
    constexpr int SS = 17;
    __shared__ T S[4][16 * SS];
    __shared__ T X[4][16 * SS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ti = blockIdx.x * 4 + w;
    if (ti >= ntask) return;
    const DiagTask t = tasks[ti];
    const int nb = t.nb;
    T *Sw = S[w], *Xw = X[w];
    if (lane < 16) {
        T v[16];
#pragma unroll
        for (int c = 0; c < 16; c++)
            v[c] = (lane < nb && c <= lane) ? W[t.off + lane + (long long)c * t.ld] : (T)0;
#pragma unroll
        for (int c = 0; c < 16; c++) Sw[lane * SS + c] = (lane >= nb && lane == c) ? (T)1 : v[c];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    bool brk = false;
    const int badj = potrf16_wave<T, SAFE>(Sw, Xw, SS, 0, nb, lane, adiag, t.col0, rho, &brk);
    if (SAFE && brk) {
        // breakdown: leave the assembled block in W for the FP64 redo
        if (lane == 0) redo_push(rd, t);
        return;
    }
    if (lane == 0 && badj < nb) atomicMin(fail, t.col0 + badj);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (lane < nb) {
#pragma unroll
        for (int c = 0; c < 16; c++) {
            if (c < nb) {
                if (c <= lane) W[t.off + lane + (long long)c * t.ld] = Sw[lane * SS + c];
                W[t.winv + lane + (long long)c * nb] = (c <= lane) ? Xw[lane * SS + c] : (T)0;
            }
        }
    }
}

// Diagonal blocks of at most TINY_DIAG columns (the bottom levels of a forest:
// mostly one to five columns): one lane factors and inverts a whole block, with
// potrf16_wave's arithmetic: pivot d_jj (when SAFE, a pivot at or below
// rho a_jj sends the block to the FP64 redo queue), a non-positive pivot
// reported and replaced by 1, r_j = 1/sqrt(d_jj), L_jj = d_jj r_j,
// L_ij = a_ij r_j, the inverse by forward substitution with r_j.
constexpr int TINY_DIAG = 6;

template <typename T, bool SAFE>
__global__ __launch_bounds__(256) void k_potrf_tiny(const DiagTask *__restrict__ tasks, int ntask,
                                                    T *__restrict__ W, int *__restrict__ fail,
                                                    const double *__restrict__ adiag, double rho, Redo rd) {
    const int ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= ntask) return;
    const DiagTask t = tasks[ti];
    const int nb = t.nb;
    T a[TINY_DIAG][TINY_DIAG], x[TINY_DIAG][TINY_DIAG], rinv[TINY_DIAG];
#pragma unroll
    for (int i = 0; i < TINY_DIAG; i++)
#pragma unroll
        for (int c = 0; c <= i; c++)
            a[i][c] = i < nb ? W[t.off + i + (long long)c * t.ld] : (i == c ? (T)1 : (T)0);
    bool low_any = false;
    int badj = TINY_DIAG;
#pragma unroll
    for (int j = 0; j < TINY_DIAG; j++) {
        const bool real = j < nb;
        T djj = a[j][j];
        if (SAFE) low_any = low_any || (real && djj <= (T)(rho * adiag[t.col0 + j]));
        const bool bad = !(djj > (T)0);
        badj = (bad && real && badj == TINY_DIAG) ? j : badj;
        djj = bad ? (T)1 : djj;
        const T rj = rsqrt_nr(djj);
        rinv[j] = rj;
        a[j][j] = djj * rj;
#pragma unroll
        for (int i = j + 1; i < TINY_DIAG; i++) a[i][j] *= rj;
#pragma unroll
        for (int c = j + 1; c < TINY_DIAG; c++)
#pragma unroll
            for (int i = c; i < TINY_DIAG; i++) a[i][c] -= a[i][j] * a[c][j];
    }
    if (SAFE && low_any) {
        redo_push(rd, t);            // breakdown: the FP64 redo refactors the assembled block
        return;
    }
    if (badj < nb) atomicMin(fail, t.col0 + badj);
#pragma unroll
    for (int c = 0; c < TINY_DIAG; c++)
#pragma unroll
        for (int i = c; i < TINY_DIAG; i++) {
            T sacc = (i == c) ? (T)1 : (T)0;
#pragma unroll
            for (int k = c; k < i; k++) sacc -= a[i][k] * x[k][c];
            x[i][c] = sacc * rinv[i];
        }
#pragma unroll
    for (int c = 0; c < TINY_DIAG; c++)
#pragma unroll
        for (int i = 0; i < TINY_DIAG; i++)
            if (i < nb && c < nb) {
                if (c <= i) W[t.off + i + (long long)c * t.ld] = a[i][c];
                W[t.winv + i + (long long)c * nb] = (c <= i) ? x[i][c] : (T)0;
            }
}

// The FP64 redo of the blocks that the one-wave and one-lane POTRF kernels
// (k_potrf_small, k_potrf_tiny) queued on breakdown: launched after each such
// launch, so a queued block is final before the next launch reads it.  A
// workgroup takes queued blocks in turn (redo_block64); the last one to
// finish clears the queue for the next launch.  With nothing queued each
// workgroup only reads the counter.
constexpr int REDO_WG = 4;

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_redo64(Redo rd, T *__restrict__ W, int *__restrict__ fail) {
    __shared__ double A[NB * (NB + 1)];
    __shared__ double tmp[NB];
    __shared__ int sbad;
    const int tid = threadIdx.x;
    const int cnt = rd.cnt[0];
    for (int e = blockIdx.x; e < cnt; e += gridDim.x) {
        redo_block64<T>(rd.list[e], W, A, tmp, &sbad, tid, (int)blockDim.x, rd, fail, [] { __syncthreads(); });
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(rd.cnt + 2, 1) == (int)gridDim.x - 1) {     // the last workgroup: clear the queue
            rd.cnt[0] = 0;
            rd.cnt[2] = 0;
        }
    }
}

// Panel TRSM of a tiny supernode (K <= TINY_DIAG columns, at most
// TINY_TRSM_ROWS rows below): one lane forms L2 = S2 inv(L11)' row by row, in
// place, as MODE_STORE does for a tile (C[i, j] = sum_k Y[i, k] X[j, k] with X
// the inverse at x_off + j + k ldx), and for a split-FP16 factor also writes
// the value's hi/lo pair scaled by its row's exponent.
constexpr int TINY_TRSM_ROWS = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_trsm_tiny(const UpdDesc *__restrict__ descs,
                                                   const Tile *__restrict__ tiles, int ntile,
                                                   T *__restrict__ W, const int *__restrict__ Ls,
                                                   unsigned *__restrict__ Lh, const double *__restrict__ adiag) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntile) return;
    const UpdDesc d = descs[tiles[e].desc];
    T inv[TINY_DIAG][TINY_DIAG];
#pragma unroll
    for (int j = 0; j < TINY_DIAG; j++)
#pragma unroll
        for (int k = 0; k < TINY_DIAG; k++)
            inv[j][k] = (j < d.N && k < d.K) ? W[d.x_off + j + (long long)k * d.ldx] : (T)0;
    for (int i = 0; i < d.M; i++) {
        T y[TINY_DIAG];
#pragma unroll
        for (int k = 0; k < TINY_DIAG; k++) y[k] = k < d.K ? W[d.y_off + i + (long long)k * d.ldy] : (T)0;
        const int ex = Lh ? row_scale_exp(adiag, Ls[d.ls_off + i]) : 0;
#pragma unroll
        for (int j = 0; j < TINY_DIAG; j++) {
            if (j >= d.N) break;
            T c = 0;
#pragma unroll
            for (int k = 0; k < TINY_DIAG; k++) c += y[k] * inv[j][k];
            const long long q = d.c_off + (long long)j * d.ldc + i;
            W[q] = c;
            if (std::is_same<T, float>::value && Lh) Lh[q] = split_pack((float)c, ex);
        }
    }
}


// ---------------------------------------------------------------------------
// Small supernodes of one etree level, fused: a supernode of at most
// SMALL_COLS columns and SMALL_NROW rows has one 64-column step, so its whole
// numeric factorization (t_cholmod_super_numeric_worker.c:537-1190: the
// descendant updates by dsyrk/dgemm scattered through the relative map, then
// dpotrf of the diagonal block and dtrsm of the rows below) is done by one
// workgroup, the supernode held in LDS from its assembly to its last store.
// The workgroups walk the level's tasks on a grid sized to the device (a
// persistent grid: task b, then b + gridDim.x, ...).  This replaces the
// level's descendant-update launch, its POTRF launch and its TRSM launch for
// these supernodes by one launch.
//
// Per descendant pair (the UpdDesc of the update launch, MODE_SCATTER
// semantics), wave w taking pairs w, w + 4, ...: C = Y X' by 16x16x4 FP64
// MFMAs straight from W (row blocks of 16 rows of Y against the N <= 16 rows
// of X), then S[rel(i)][rel(j)] -= C(i, j) for i >= j by LDS atomics, rel the
// block rows' positions among the supernode's rows.  Factorization, by wave 0:
// right-looking over the nscol columns with lane r owning row r of the panel,
// potrf16_wave's pivot arithmetic (r_j = 1/sqrt(d_jj), L_jj = d_jj r_j,
// L_rj = a_rj r_j, a non-positive pivot reported and replaced by 1) applied
// to the rows below the diagonal block as well, which makes the TRSM part of
// the same pass.  The inverse of the diagonal block is formed as potrf16_wave
// forms it, column c by lane c.
// ---------------------------------------------------------------------------
constexpr int SMALL_COLS = 16;
constexpr int SMALL_NROW = 64;

__global__ __launch_bounds__(256) void k_small_level(const SmallTask *__restrict__ tasks, int ntask,
                                                     const UpdDesc *__restrict__ descs, double *__restrict__ W,
                                                     const int *__restrict__ Ls, int *__restrict__ fail) {
    constexpr int SD = SMALL_COLS + 1;
    __shared__ double S[SMALL_NROW * SD];
    __shared__ int rows[SMALL_NROW];
    __shared__ int rel[4][SMALL_NROW];
    __shared__ double rpiv[SMALL_COLS];        // r_j = 1/sqrt(d_jj) of each column
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int jl = lane & 15, kq = lane >> 4;
    int *rl = rel[w];
    for (int ti = blockIdx.x; ti < ntask; ti += gridDim.x) {
        const SmallTask t = tasks[ti];
        const int nsrow = t.nsrow, nscol = t.nscol;
        // the assembled block (thread e: row e % 64 of columns e / 64, e / 64 + 4, ...) and its rows
#pragma unroll
        for (int q = 0; q < SMALL_NROW * SMALL_COLS / 256; q++) {
            const int r = lane, c = w + 4 * q;
            S[r * SD + c] = (r < nsrow && c < nscol) ? W[t.lx + r + (long long)c * nsrow] : 0.0;
        }
        if (w == 0) rows[lane] = lane < nsrow ? Ls[t.psi + lane] : INT_MAX;
        __syncthreads();

#pragma unroll 1
        for (int p = t.u0 + w; p < t.u0 + t.nu; p += 4) {
            const UpdDesc u = descs[p];
            if (lane < u.M) {
                const int g = Ls[u.ls_off + lane];
                int r;
                if (g < t.k1 + nscol) {
                    r = g - t.k1;
                } else {
                    int lo = nscol, hi = nsrow - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (rows[mid] < g) lo = mid + 1; else hi = mid;
                    }
                    r = lo;
                }
                rl[lane] = r;
            }
            const int nblk = (u.M + 15) >> 4;
            f64x4 acc[4];
#pragma unroll
            for (int n = 0; n < 4; n++) acc[n] = f64x4{0, 0, 0, 0};
            // 8 columns of the descendant per pass: all 10 loads of a lane in
            // flight at once (32-bit offsets from the pair's base), then the
            // MFMAs in ascending k
            const double *xb = W + u.x_off, *yb0 = W + u.y_off;
#pragma unroll 1
            for (int k0 = 0; k0 < u.K; k0 += 8) {
                double xa[2], yb[2][4];
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const int k = k0 + 4 * c + kq;
                    const bool kin = k < u.K;
                    xa[c] = (kin && jl < u.N) ? xb[jl + k * u.ldx] : 0.0;
#pragma unroll
                    for (int n = 0; n < 4; n++)
                        yb[c][n] = (kin && 16 * n + jl < u.M) ? yb0[16 * n + jl + k * u.ldy] : 0.0;
                }
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (k0 + 4 * c >= u.K) break;
#pragma unroll
                    for (int n = 0; n < 4; n++)
                        if (n < nblk) acc[n] = MfmaOp<double>::mma(xa[c], yb[c][n], acc[n]);
                }
            }
            wave_sync_lds();                 // rel[] of every lane of the wave
            // acc[n][v] = C(i, j), i = 16 n + lane % 16, j = row(lane, v)
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const int i = 16 * n + jl;
                if (n >= nblk || i >= u.M) continue;
                const int ri = rl[i];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int j = MfmaOp<double>::row(lane, v);
                    if (j < u.N && j <= i)
                        __hip_atomic_fetch_add(&S[ri * SD + rl[j]], -acc[n][v], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            wave_sync_lds();                 // the map's reads, before the next pair's map
        }
        __syncthreads();

        if (w == 0) {
            // the panel, right-looking, lane r owning row r
            double a[SMALL_COLS];
#pragma unroll
            for (int c = 0; c < SMALL_COLS; c++) a[c] = c < nscol ? S[lane * SD + c] : 0.0;
            int badj = SMALL_COLS;
#pragma unroll
            for (int j = 0; j < SMALL_COLS; j++) {
                if (j >= nscol) continue;
                double djj = lane_bcast(a[j], j);
                const bool bad = !(djj > 0.0);
                badj = (bad && badj == SMALL_COLS) ? j : badj;
                djj = bad ? 1.0 : djj;
                const double rj = rsqrt_nr(djj);
                if (lane == j) rpiv[j] = rj;
                const double l = (lane > j) ? a[j] * rj : ((lane == j) ? djj * rj : 0.0);
                a[j] = l;
#pragma unroll
                for (int c = j + 1; c < SMALL_COLS; c++)
                    if (c < nscol) a[c] -= l * lane_bcast(l, c);
            }
            if (lane == 0 && badj < nscol) atomicMin(fail, t.k1 + badj);
            if (lane < nsrow) {
#pragma unroll
                for (int c = 0; c < SMALL_COLS; c++)
                    if (c < nscol && c <= lane) W[t.lx + lane + (long long)c * nsrow] = a[c];
            }
            // the inverse of the diagonal block: lane c forms column c
            if (lane < nscol) {
#pragma unroll
                for (int c = 0; c < SMALL_COLS; c++) S[lane * SD + c] = a[c];
            }
            wave_sync_lds();
            // x overwrites the partial sums in place: x_k = sm_k r_k once sm_k is final
            double x[SMALL_COLS];
#pragma unroll
            for (int i = 0; i < SMALL_COLS; i++) x[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < SMALL_COLS; k++) {
                if (k >= nscol) break;
                x[k] = x[k] * rpiv[k];
#pragma unroll
                for (int i = k + 1; i < SMALL_COLS; i++)
                    if (i < nscol) x[i] -= S[i * SD + k] * x[k];
            }
            if (lane < nscol) {
#pragma unroll
                for (int i = 0; i < SMALL_COLS; i++)
                    if (i < nscol) W[t.winv + i + (long long)lane * nscol] = (i >= lane) ? x[i] : 0.0;
            }
        }
        __syncthreads();                     // this task's LDS reads, before the next task's stores
    }
}

}  // namespace mxc
