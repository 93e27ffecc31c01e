// symbolic_internal.h -- layout of the CPU symbolic factor (shared by the C
// symbolic code and the C++/HIP plan builder).
#ifndef MXC_SYMBOLIC_INTERNAL_H
#define MXC_SYMBOLIC_INTERNAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct mxc_symbolic {
    int64_t n;
    int64_t nsuper;
    int64_t ssize;
    int64_t xsize;
    int64_t maxcsize;
    int64_t maxesize;
    int ordering;
    double fl;
    double lnz;
    // ordering strategy record (Common->selected, method[0..2].fl/lnz, anz):
    // slot 0 GIVEN, 1 AMD (AMD's own predictions), 2 METIS; -1 = not tried
    int selected;
    double method_fl[3], method_lnz[3], anz;
    int64_t *Perm;      // n: Perm[k] = original index of pivot k
    int64_t *ColCount;  // n (NULL after import)
    int64_t *Parent;    // n, etree of the permuted matrix (NULL after import)
    int64_t *Super;     // nsuper+1
    int64_t *Lpi;       // nsuper+1
    int64_t *Lpx;       // nsuper+1
    int64_t *Ls;        // ssize
};

// Columns of the supernodal L reached from C[0..blen), in lsolve_pattern order.
int mxc_reach(int64_t n, int64_t nsuper, const int64_t *Super, const int64_t *Lpi, const int64_t *Ls,
              const int64_t *C, int64_t blen, int64_t *Yset, int64_t *ylen);

#ifdef __cplusplus
}
#endif
#endif
