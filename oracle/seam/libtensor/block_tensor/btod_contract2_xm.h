// Test infrastructure (checker only, never the product).
//
// Seam header for ONE translation unit of the reference build:
// libtensor/expr/btensor/impl/eval_btensor_double_contract.C compiled with
// -DWITH_LIBXM and this directory first on the include path.  Its
// alternate-backend branch (eval_btensor_double_contract.C:205-216, taken
// when eval_btensor<double>::use_libxm(true) is set) then builds the MI355X
// adapter where it would build libxm's btod_contract2_xm; with
// use_libxm(false) the same evaluator builds libtensor's btod_contract2.  In
// an integrated libtensor this is the three-line branch INTEGRATION.md shows
// (the adapter selected by its own switch); here it lets the unmodified
// reference sources drive the adapter.
#ifndef LIBTENSOR_BTOD_CONTRACT2_XM_SEAM_H
#define LIBTENSOR_BTOD_CONTRACT2_XM_SEAM_H

#include "../../../../libtensor_amd/integration/btod_contract2_amd.h"

namespace libtensor {

template<size_t N, size_t M, size_t K>
class btod_contract2_xm : public btod_contract2_amd<N, M, K> {
public:
    btod_contract2_xm(const contraction2<N, M, K> &contr,
        block_tensor_rd_i<N + K, double> &bta, double ka,
        block_tensor_rd_i<M + K, double> &btb, double kb, double kc) :
        btod_contract2_amd<N, M, K>(contr, bta, ka, btb, kb, kc) { }
};

} // namespace libtensor

#endif // LIBTENSOR_BTOD_CONTRACT2_XM_SEAM_H
