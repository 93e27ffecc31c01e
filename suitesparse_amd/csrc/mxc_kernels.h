// mxc_kernels.h -- gfx950 kernels of the supernodal numeric factorization and
// solve.  Templated on the working type T (double: FP64 MFMA, float: FP32 MFMA).
//
// Storage: one device buffer W holds L->x (xsize entries, CHOLMOD's layout:
// supernode s is a column-major nsrow x nscol block at Lpx[s], ld = nsrow)
// followed by the inverses of the 64-column diagonal blocks of every
// supernode (block (s,t) of width nb stored column-major with ld = nb).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mxc {

constexpr int TILE = 64;     // C tile edge (rows and columns)
constexpr int BK = 16;       // K chunk staged through LDS
constexpr int NB = 64;       // diagonal block width of the blocked POTRF
constexpr int DS = NB + 1;   // LDS row stride of a 64x64 diagonal block
constexpr int LDS_PAD = 4;

// One dense update C = Yp * Xp' (C is M x N) applied to a target.
//   X side (C columns): row j of Xp is W[x_off + j + k*ldx], k < K
//   Y side (C rows):    row i of Yp is W[y_off + i + k*ldy], k < K
// MODE_SCATTER: W[c_off + relc(j)*ldc + relr(i)] -= C(i,j) for i >= j, with the
//   relative map of a descendant d into ancestor s (t_* describe s's rows):
//   C row i is global row Ls[ls_off+i]; it lands on local row g-base if
//   g < base+t_nscol (one of s's own columns) else at its position in
//   Ls[t_psi+t_nscol .. t_psi+t_nsrow); C column j lands on local column g-base.
// MODE_TRAIL:   W[c_off + j*ldc + i] -= C(i,j) for i >= j (in-place trailing update)
// MODE_STORE:   W[c_off + j*ldc + i]  = C(i,j)           (TRSM by an inverse block)
// Block C(i, j) = sum_k Y[i][k] X[j][k], i < M, j < N (X rows are the first N
// rows of the Y side); ls_off: Ls position of Y row 0 (its global row index).
struct UpdDesc {
    long long x_off, y_off, c_off;
    int ldx, ldy, ldc;
    int K, M, N;
    int ls_off, base;
    int t_psi, t_nscol, t_nsrow;
    int pd;            // trailing blocks: diagonal task factored after tile (0, 0), or -1
    int fold;          // trailing blocks with pd >= 0 (FP64, 64-wide tiles): the tiles of
                       // column 0 below row 64 also apply the next step's TRSM (by the
                       // inverse of diagonal task pd, once tile (0, 0) has published it)
};

struct Tile {
    int desc;
    int ti;
    int tj;
};

// Diagonal block (s, t): rows/cols [c0, c0+nb) of supernode s.
struct DiagTask {
    long long off;     // W offset of element (c0, c0)
    long long winv;    // W offset of the inverse block (ld nb)
    int ld;            // nsrow of s
    int nb;
    int col0;          // global column index k1 + c0 (for the failing-column report)
    int pad;
};

// A small supernode factored whole by one wave (k_small_level): its
// descendant updates are descs[u0 .. u0 + nu), in descendant order.
struct SmallTask {
    long long lx;      // W offset of the supernode (Lpx[s])
    long long winv;    // W offset of the inverse of its diagonal block (ld nscol)
    int nsrow, nscol;
    int k1, psi;       // first column, Ls offset of its rows
    int u0, nu;
};

// Solve work item: column block (s, t) and a row range [r0, r1) of s.
// One workgroup of a supernodal solve step (forward or backward sweep).
//   diag: block t of supernode s, columns [c0, c0 + nb); r0 = width of the
//         neighbouring block folded in (forward: block t-1, backward: the rows
//         of block t+1; 0 = none); backward: reduces partial slots [p0, p0+np).
//   off:  block columns [c0, c0 + nb) against rows [r0, r1) of the supernode
//         (rows >= nscol map through Ls); backward: writes partial slot p0.
struct SweepTask {
    long long lx;      // W offset of supernode s (Lpx[s])
    long long winv;    // diag: inverse of the block (nb x nb, ld nb)
    int nsrow, nscol;
    int k1, psi;
    int c0, nb;
    int r0, r1;
    int p0, np;
    int pad[2];
    // chained backward sweep: the slots [q0, q0 + nq) of the rows below block t,
    // complete once counter bc reaches bt
    int q0, nq, bc, bt;
};

enum { MODE_SCATTER = 0, MODE_TRAIL = 1, MODE_STORE = 2 };

}  // namespace mxc
