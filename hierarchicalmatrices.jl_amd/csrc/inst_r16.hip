// inst_r16.hip -- kernels for ranks in bucket 16 (r <= 16).
#include "kernels_impl.h"

hipError_t hmx_matvec_r16(const DevPlan &p, const MvArgs &a, bool tr, int mode, int kb,
                          hipStream_t s, const hipEvent_t *ev) {
    return hmxk::matvec_r<double, 16>(p, a, tr, mode, kb, s, ev);
}

hipError_t hmx_grad_r16(const DevPlan &p, const MvArgs &a, bool tr, double *y2, int64_t incy2,
                        const double *sv, int64_t incs, hipStream_t s) {
    return hmxk::grad_r<double, 16>(p, a, tr, y2, incy2, sv, incs, s);
}

hipError_t hmx_gen_basis_r16(const DevPlan &p, hipStream_t s) { return hmxk::gen_basis_r<double, 16>(p, s); }

hipError_t hmx_matvec_m_r16(const DevPlan &p, const MvArgs &a, int mode, int prec, int nq,
                            hipStream_t s) {
    return hmxk::matvec_m_r<16>(p, a, mode, prec, nq, s);
}
