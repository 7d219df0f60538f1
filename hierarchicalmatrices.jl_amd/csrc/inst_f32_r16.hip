// inst_f32_r16.hip -- kernels of FP32 plans for ranks in bucket 16 (r <= 16).
#include "kernels_impl.h"

hipError_t hmx_matvec_f32_r16(const DevPlanF &p, const MvArgsF &a, bool tr, int mode, int kb,
                              hipStream_t s, const hipEvent_t *ev) {
    return hmxk::matvec_r<float, 16>(p, a, tr, mode, kb, s, ev);
}

hipError_t hmx_grad_f32_r16(const DevPlanF &p, const MvArgsF &a, bool tr, float *y2,
                            int64_t incy2, const float *sv, int64_t incs, hipStream_t s) {
    return hmxk::grad_r<float, 16>(p, a, tr, y2, incy2, sv, incs, s);
}

hipError_t hmx_gen_basis_f32_r16(const DevPlanF &p, hipStream_t s) {
    return hmxk::gen_basis_r<float, 16>(p, s);
}
