// bmg_rb.h -- launchers of the RB block kernels (bmg_rb.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "bmg_chain.h"

namespace bmg {
int launch_residuals(const double* y, uint32_t n, uint32_t n_pad, const ChainArrays* arrays,
                     const int* list, int nv, double* r, double* rsum, hipStream_t s);
int launch_rb_finalize(const DevData* data, const ChainConst* consts, const ChainScal* scal,
                       const ChainArrays* arrays, const int* list, int nv, int kmax, uint64_t p,
                       const double* xa, const double* xh, const double* rsum, hipStream_t s);
int launch_rb_update(const DevData* data, const ChainConst* consts, const ChainScal* scal,
                     const ChainArrays* arrays, const int* list, int nv, uint64_t p, hipStream_t s);
// chain `chain` takes the table p (device, np values in [0, 1]) as its
// p_proposal, with q_add, q_rem and the tree leaves following; the caller
// rebuilds the tree levels
int launch_set_proposal(const DevData* data, const ChainConst* consts, const ChainArrays* arrays, int chain,
                        const double* p, uint64_t np, hipStream_t s);
int launch_tree_build(const DevData* data, const ChainConst* consts, ChainScal* scal,
                      const ChainArrays* arrays, const int* list, int nv, const uint64_t* lev_len,
                      int top, int only_rebuild, int clear_pending, hipStream_t s);
}  // namespace bmg
