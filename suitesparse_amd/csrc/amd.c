// amd.c -- approximate minimum degree ordering (placeholder until the
// restatement of AMD lands).
#include "../../include/cholmod_mi355x.h"

int mxc_amd_order(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *Perm) {
    (void)n; (void)Ap; (void)Ai; (void)Perm;
    return MXC_NOT_INSTALLED;
}
